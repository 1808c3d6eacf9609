!=================================================================
! Fortran drop-in modules over the C-ABI of librdsphp
! (include/redondito.h).  They keep the names the reference's callers
! USE:
!   gdevice  -- device_init, device_sync, gdev_alloc, gdev_free,
!               gdev_lookup, gdev_update_to/from, gdev_copy, gdev_present,
!               gdev_active (src/memmgt/gdevice_mod.f90)
!   fcgram   -- TYPE(FCCOORD), TYPE(FFTPLAN), TYPE(FCPLAN),
!               fcgram_create_plan, fcgram_destroy_plan,
!               load_dirichlet_tables, load_neumann_tables,
!               neumann_reconstruct, robin_reconstruct (src/sphp/fcgram_mod.f90,
!               src/sphp-gpu/sphp_mod.F90:43-54)
!   fft      -- the global plan planfc (src/pseudo/pseudospec_mod.f90)
!   grid, kes, boxsize -- grid_init, box_init and their grids
!               (src/pseudo/pseudospec_mod.f90)
! Field arrays stay the caller's host arrays; their device copies are
! registered with gdev_alloc and found by host address, as in the
! reference's GPU build.  Errors print and STOP, as the reference does.
!=================================================================

      MODULE rd_capi
      USE iso_c_binding
      IMPLICIT NONE
      INTERFACE
        FUNCTION rd_error_string(code) BIND(C,NAME='rd_error_string')
          IMPORT :: C_INT, C_PTR
          INTEGER(C_INT), VALUE :: code
          TYPE(C_PTR) :: rd_error_string
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_load_dirichlet_tables(tdir,C,d,dir) &
                BIND(C,NAME='rd_load_dirichlet_tables')
          IMPORT :: C_INT, C_CHAR, C_DOUBLE
          CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: tdir
          INTEGER(C_INT), VALUE :: C, d
          REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: dir
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_plan_create(plan,lmax,nlat,nphi,nz,C,d,dz,rint,dir) &
                BIND(C,NAME='rd_plan_create')
          IMPORT :: C_INT, C_PTR, C_DOUBLE
          TYPE(C_PTR), INTENT(OUT) :: plan
          INTEGER(C_INT), VALUE :: lmax, nlat, nphi, nz, C, d
          REAL(C_DOUBLE), VALUE :: dz, rint
          TYPE(C_PTR), VALUE :: dir
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_plan_destroy(plan) BIND(C,NAME='rd_plan_destroy')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_plan_sync(plan) BIND(C,NAME='rd_plan_sync')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_plan_grid(plan,ct,st,wt,z,kz,ell,emm) &
                BIND(C,NAME='rd_plan_grid')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, ct, st, wt, z, kz, ell, emm
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_sht_analysis(plan,nplanes,a,b) BIND(C,NAME='rd_sht_analysis')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, a, b
          INTEGER(C_INT), VALUE :: nplanes
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_sht_synthesis(plan,nplanes,a,b) BIND(C,NAME='rd_sht_synthesis')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, a, b
          INTEGER(C_INT), VALUE :: nplanes
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_vsht_analysis(plan,nplanes,vr,vt,vp,q,s,t) &
                BIND(C,NAME='rd_vsht_analysis')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, vr, vt, vp, q, s, t
          INTEGER(C_INT), VALUE :: nplanes
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_vsht_synthesis(plan,nplanes,q,s,t,vr,vt,vp) &
                BIND(C,NAME='rd_vsht_synthesis')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, q, s, t, vr, vt, vp
          INTEGER(C_INT), VALUE :: nplanes
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_z_forward(plan,a) BIND(C,NAME='rd_z_forward')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, a
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_z_inverse(plan,a) BIND(C,NAME='rd_z_inverse')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, a
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_3d_forward(plan,a,b) BIND(C,NAME='rd_3d_forward')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, a, b
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_3d_inverse(plan,a,b) BIND(C,NAME='rd_3d_inverse')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, a, b
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_derivr(plan,p,esp,a,b) BIND(C,NAME='rd_derivr')
          IMPORT :: C_INT, C_PTR, C_CHAR
          TYPE(C_PTR), VALUE :: plan, a, b
          INTEGER(C_INT), VALUE :: p
          CHARACTER(KIND=C_CHAR), VALUE :: esp
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_laplang(plan,a,b) BIND(C,NAME='rd_laplang')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, a, b
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_multr(plan,p,a,b) BIND(C,NAME='rd_multr')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, a, b
          INTEGER(C_INT), VALUE :: p
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_dlr(plan,a,b) BIND(C,NAME='rd_dlr')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, a, b
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_saxpby_c(plan,z,x,a,y,b) BIND(C,NAME='rd_saxpby_c')
          IMPORT :: C_INT, C_PTR, C_DOUBLE
          TYPE(C_PTR), VALUE :: plan, z, x, y
          REAL(C_DOUBLE), VALUE :: a, b
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_saxpby_r(plan,nplanes,z,x,a,y,b) BIND(C,NAME='rd_saxpby_r')
          IMPORT :: C_INT, C_PTR, C_DOUBLE
          TYPE(C_PTR), VALUE :: plan, z, x, y
          INTEGER(C_INT), VALUE :: nplanes
          REAL(C_DOUBLE), VALUE :: a, b
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_prodvec(plan,nplanes,ar,at,ap,br,bt,bp,cr,ct,cp) &
                BIND(C,NAME='rd_prodvec')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, ar, at, ap, br, bt, bp, cr, ct, cp
          INTEGER(C_INT), VALUE :: nplanes
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_fc_filter(plan,esp,a) BIND(C,NAME='rd_fc_filter')
          IMPORT :: C_INT, C_PTR, C_CHAR
          TYPE(C_PTR), VALUE :: plan, a
          CHARACTER(KIND=C_CHAR), VALUE :: esp
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_load_neumann_tables(tdir,d,ord,dx,neu) &
                BIND(C,NAME='rd_load_neumann_tables')
          IMPORT :: C_INT, C_CHAR, C_DOUBLE
          CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: tdir
          INTEGER(C_INT), VALUE :: d, ord
          REAL(C_DOUBLE), VALUE :: dx
          REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: neu
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_neumann_reconstruct(plan,a,boun,ord,neu) &
                BIND(C,NAME='rd_neumann_reconstruct')
          IMPORT :: C_INT, C_PTR, C_DOUBLE
          TYPE(C_PTR), VALUE :: plan, a
          INTEGER(C_INT), VALUE :: boun, ord
          REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: neu
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_robin_reconstruct(plan,a,boun,neu,coef) &
                BIND(C,NAME='rd_robin_reconstruct')
          IMPORT :: C_INT, C_PTR, C_DOUBLE
          TYPE(C_PTR), VALUE :: plan, a, coef
          INTEGER(C_INT), VALUE :: boun
          REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: neu
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_redist_k_to_lm(plan,nplanes,a,b) BIND(C,NAME='rd_redist_k_to_lm')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, a, b
          INTEGER(C_INT), VALUE :: nplanes
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_redist_lm_to_k(plan,nplanes,a,b) BIND(C,NAME='rd_redist_lm_to_k')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, a, b
          INTEGER(C_INT), VALUE :: nplanes
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_dev_index(h) BIND(C,NAME='rd_dev_index')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: h
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_copy(plan,a,b) BIND(C,NAME='rd_copy')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, a, b
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_dirichlet(plan,a) BIND(C,NAME='rd_dirichlet')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, a
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_nonlinear_uxw(plan,prec,w,z,nq,ns,nt) &
                BIND(C,NAME='rd_nonlinear_uxw')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, w, z, nq, ns, nt
          INTEGER(C_INT), VALUE :: prec
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_nonlinear_coupled(plan,prec,w,z,pb,tb,t,nq,ns,nt,mq,ms,mt,nh) &
                BIND(C,NAME='rd_nonlinear_coupled')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, w, z, pb, tb, t, nq, ns, nt, mq, ms, mt, nh
          INTEGER(C_INT), VALUE :: prec
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_derivr_mat(plan,p,d) BIND(C,NAME='rd_derivr_mat')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, d
          INTEGER(C_INT), VALUE :: p
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_bc_setup(plan,bcint,bcext) BIND(C,NAME='rd_bc_setup')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan
          INTEGER(C_INT), VALUE :: bcint, bcext
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_invdl(plan,a,b) BIND(C,NAME='rd_invdl')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, a, b
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_v_imposebc(plan,tor,phi,pol) BIND(C,NAME='rd_v_imposebc')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, tor, phi, pol
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_v_imposebc_fast(plan,tor,phi,pol) BIND(C,NAME='rd_v_imposebc_fast')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: plan, tor, phi, pol
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_device_init(rank) BIND(C,NAME='rd_device_init')
          IMPORT :: C_INT
          INTEGER(C_INT), VALUE :: rank
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_device_sync() BIND(C,NAME='rd_device_sync')
          IMPORT :: C_INT
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_dev_alloc(h,nb) BIND(C,NAME='rd_dev_alloc')
          IMPORT :: C_INT, C_PTR, C_SIZE_T
          TYPE(C_PTR), VALUE :: h
          INTEGER(C_SIZE_T), VALUE :: nb
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_dev_free(h) BIND(C,NAME='rd_dev_free')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: h
        END FUNCTION
        TYPE(C_PTR) FUNCTION rd_dev_lookup(h) BIND(C,NAME='rd_dev_lookup')
          IMPORT :: C_PTR
          TYPE(C_PTR), VALUE :: h
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_dev_present(h) BIND(C,NAME='rd_dev_present')
          IMPORT :: C_INT, C_PTR
          TYPE(C_PTR), VALUE :: h
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_dev_update_to(h,nb) BIND(C,NAME='rd_dev_update_to')
          IMPORT :: C_INT, C_PTR, C_SIZE_T
          TYPE(C_PTR), VALUE :: h
          INTEGER(C_SIZE_T), VALUE :: nb
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_dev_update_from(h,nb) BIND(C,NAME='rd_dev_update_from')
          IMPORT :: C_INT, C_PTR, C_SIZE_T
          TYPE(C_PTR), VALUE :: h
          INTEGER(C_SIZE_T), VALUE :: nb
        END FUNCTION
        INTEGER(C_INT) FUNCTION rd_dev_copy(hd,hs,nb) BIND(C,NAME='rd_dev_copy')
          IMPORT :: C_INT, C_PTR, C_SIZE_T
          TYPE(C_PTR), VALUE :: hd, hs
          INTEGER(C_SIZE_T), VALUE :: nb
        END FUNCTION
      END INTERFACE

      CONTAINS

! Prints the library's message for a failed call and stops, like the
! reference's PRINT + STOP error handling.
      SUBROUTINE rd_check(rc,what)
      INTEGER(C_INT), INTENT(IN) :: rc
      CHARACTER(len=*), INTENT(IN) :: what
      CHARACTER(len=128), POINTER :: msg
      CHARACTER(len=128) :: txt
      INTEGER :: i
      IF (rc .eq. 0) RETURN
      CALL C_F_POINTER(rd_error_string(rc),msg)
      txt = msg
      i = INDEX(txt,C_NULL_CHAR)
      IF (i .gt. 0) txt(i:) = ' '
      PRINT *, '[ERROR] ', what, ': ', TRIM(txt)
      STOP
      END SUBROUTINE rd_check

! Device address of a registered host array (stops when it has none,
! as the reference does for an unregistered array).
      FUNCTION rd_dptr(h,what) RESULT(d)
      TYPE(C_PTR), INTENT(IN) :: h
      CHARACTER(len=*), INTENT(IN) :: what
      TYPE(C_PTR) :: d
      d = rd_dev_lookup(h)
      IF (.NOT. C_ASSOCIATED(d)) THEN
         PRINT *, '[ERROR] ', what, ': array has no device copy (gdev_alloc)'
         STOP
      ENDIF
      END FUNCTION rd_dptr

      END MODULE rd_capi

!=================================================================
      MODULE gdevice
      USE iso_c_binding
      USE rd_capi
      IMPLICIT NONE
! Kept for source compatibility: the device path is the only path.
      LOGICAL, SAVE :: gdev_active = .TRUE.
      CONTAINS

      SUBROUTINE device_init(myrank)
      INTEGER, INTENT(IN) :: myrank
      CALL rd_check(rd_device_init(INT(myrank,C_INT)),'device_init')
      END SUBROUTINE device_init

      SUBROUTINE device_sync()
      CALL rd_check(rd_device_sync(),'device_sync')
      END SUBROUTINE device_sync

      SUBROUTINE gdev_alloc(hp,nbytes)
      TYPE(C_PTR), INTENT(IN) :: hp
      INTEGER(C_SIZE_T), INTENT(IN) :: nbytes
      CALL rd_check(rd_dev_alloc(hp,nbytes),'gdev_alloc')
      END SUBROUTINE gdev_alloc

      SUBROUTINE gdev_free(hp)
      TYPE(C_PTR), INTENT(IN) :: hp
      CALL rd_check(rd_dev_free(hp),'gdev_free')
      END SUBROUTINE gdev_free

      SUBROUTINE gdev_update_to(hp,nbytes)
      TYPE(C_PTR), INTENT(IN) :: hp
      INTEGER(C_SIZE_T), INTENT(IN) :: nbytes
      CALL rd_check(rd_dev_update_to(hp,nbytes),'gdev_update_to')
      END SUBROUTINE gdev_update_to

      SUBROUTINE gdev_update_from(hp,nbytes)
      TYPE(C_PTR), INTENT(IN) :: hp
      INTEGER(C_SIZE_T), INTENT(IN) :: nbytes
      CALL rd_check(rd_dev_update_from(hp,nbytes),'gdev_update_from')
      END SUBROUTINE gdev_update_from

      SUBROUTINE gdev_copy(hdst,hsrc,nbytes)
      TYPE(C_PTR), INTENT(IN) :: hdst, hsrc
      INTEGER(C_SIZE_T), INTENT(IN) :: nbytes
      CALL rd_check(rd_dev_copy(hdst,hsrc,nbytes),'gdev_copy')
      END SUBROUTINE gdev_copy

! Index of the array registered at host address hp, 0 if it has no device
! copy (gdevice_mod.f90:220-235).
      INTEGER FUNCTION gdev_lookup(hp)
      TYPE(C_PTR), INTENT(IN) :: hp
      gdev_lookup = INT(rd_dev_index(hp))
      END FUNCTION gdev_lookup

      LOGICAL FUNCTION gdev_present(hp)
      TYPE(C_PTR), INTENT(IN) :: hp
      gdev_present = rd_dev_present(hp) .ne. 0
      END FUNCTION gdev_present

      END MODULE gdevice

!=================================================================
      MODULE fcgram
      USE iso_c_binding
      USE rd_capi
      IMPLICIT NONE

      TYPE FCCOORD
         DOUBLE PRECISION, ALLOCATABLE, DIMENSION(:,:) :: dir
         DOUBLE PRECISION, ALLOCATABLE, DIMENSION(:)   :: neu,neu2
         DOUBLE PRECISION :: dx
         INTEGER          :: N,C,d
      END TYPE FCCOORD

! The reference's FFTPLAN (src/sphp-gpu/sphp_mod.F90:43-54).  The scratch
! arrays live in the library plan on the device, so the array pointers stay
! disassociated; planrxy holds the library plan in place of the SHTns
! handle.  The MPI block types of the multi-rank exchange have no
! counterpart on one GPU.
      TYPE FFTPLAN
         DOUBLE COMPLEX, DIMENSION(:,:), POINTER     :: ccarr => NULL()
         DOUBLE COMPLEX, DIMENSION(:,:), POINTER     :: carr1 => NULL()
         DOUBLE COMPLEX, DIMENSION(:,:), POINTER     :: carr2 => NULL()
         DOUBLE COMPLEX, DIMENSION(:,:), POINTER     :: carr3 => NULL()
         DOUBLE PRECISION, DIMENSION(:,:,:), POINTER :: rarr => NULL()
         TYPE(C_PTR)     :: planrxy = C_NULL_PTR
         INTEGER(C_LONG) :: plancz = 0
         INTEGER :: nlat,nphi,nz,nlm,lmax
      END TYPE FFTPLAN

      TYPE FCPLAN
         TYPE(FCCOORD)      :: x,y,z
         TYPE(FFTPLAN)      :: planrc, plancr
         CHARACTER(len=128) :: tdir
         TYPE(C_PTR)        :: h = C_NULL_PTR
      END TYPE FCPLAN

      CONTAINS

!-----------------------------------------------------------------
! fcgram_create_plan(this,N,C,o,dx,tdir,fftflags,lmax), the
! reference's constructor (src/sphp/fcgram_mod.f90:49-150): only z may
! be non-periodic; the z tables are read from tdir.  rint is the first
! radius (z(1)); it only enters the nonlinear term (1/r factors) and
! defaults to 0.
      SUBROUTINE fcgram_create_plan(this,N,C,o,dx,tdir,fftflags,lmax,rint)
      TYPE(FCPLAN), INTENT(OUT)    :: this
      INTEGER, INTENT(IN)          :: N(3), C(3), o(3), fftflags
      DOUBLE PRECISION, INTENT(IN) :: dx(3)
      CHARACTER(len=*), INTENT(IN) :: tdir
      INTEGER, INTENT(IN)          :: lmax
      DOUBLE PRECISION, INTENT(IN), OPTIONAL :: rint
      DOUBLE PRECISION, TARGET, ALLOCATABLE :: dv(:)
      DOUBLE PRECISION :: r0
      INTEGER :: nlm

      IF (C(1).ne.0 .OR. o(1).ne.0 .OR. C(2).ne.0 .OR. o(2).ne.0) THEN
         PRINT*, "[ERROR] FC-Gram continuation is only ", &
            "supported in z (Cx = Cy = 0 and ox = oy = 0). Aborting..."
         STOP
      ENDIF
      IF ((C(3).eq.0) .neqv. (o(3).eq.0)) THEN
         PRINT*, "[ERROR] Mismatch in continuation", &
            " or matching points in z direction. Aborting..."
         STOP
      ENDIF
      this%x%N = N(1); this%y%N = N(2); this%z%N = N(3)
      this%x%C = 0;    this%y%C = 0;    this%z%C = C(3)
      this%x%d = 0;    this%y%d = 0;    this%z%d = o(3)
      this%x%dx = dx(1); this%y%dx = dx(2); this%z%dx = dx(3)
      this%tdir = tdir
      nlm = (lmax+1)*(lmax+2)/2
      this%planrc%nlat = N(1); this%planrc%nphi = N(2); this%planrc%nz = N(3)
      this%planrc%nlm = nlm;   this%planrc%lmax = lmax
      r0 = 0.0D0
      IF (PRESENT(rint)) r0 = rint
      IF (C(3) .gt. 0) THEN
         CALL load_dirichlet_tables(this%z,tdir)
         ALLOCATE(dv(C(3)*o(3)))
         dv = RESHAPE(this%z%dir,(/C(3)*o(3)/))
         CALL rd_check(rd_plan_create(this%h,INT(lmax,C_INT),INT(N(1),C_INT), &
              INT(N(2),C_INT),INT(N(3),C_INT),INT(C(3),C_INT),INT(o(3),C_INT), &
              REAL(dx(3),C_DOUBLE),REAL(r0,C_DOUBLE),C_LOC(dv)),'fcgram_create_plan')
         DEALLOCATE(dv)
      ELSE
         CALL rd_check(rd_plan_create(this%h,INT(lmax,C_INT),INT(N(1),C_INT), &
              INT(N(2),C_INT),INT(N(3),C_INT),0_C_INT,0_C_INT, &
              REAL(dx(3),C_DOUBLE),REAL(r0,C_DOUBLE),C_NULL_PTR),'fcgram_create_plan')
      ENDIF
      this%planrc%planrxy = this%h
      this%plancr = this%planrc
      END SUBROUTINE fcgram_create_plan

      SUBROUTINE fcgram_destroy_plan(this)
      TYPE(FCPLAN), INTENT(INOUT) :: this
      IF (ALLOCATED(this%z%dir)) DEALLOCATE(this%z%dir)
      IF (ALLOCATED(this%z%neu)) DEALLOCATE(this%z%neu)
      IF (ALLOCATED(this%z%neu2)) DEALLOCATE(this%z%neu2)
      IF (C_ASSOCIATED(this%h)) CALL rd_check(rd_plan_destroy(this%h),'fcgram_destroy_plan')
      this%h = C_NULL_PTR
      END SUBROUTINE fcgram_destroy_plan

! dir = A.Q^T from A{C}-{d}.dat and Q{d}.dat (fcgram_mod.f90:186-263).
      SUBROUTINE load_dirichlet_tables(coord,tdir)
      TYPE(FCCOORD), INTENT(INOUT) :: coord
      CHARACTER(len=*), INTENT(IN) :: tdir
      DOUBLE PRECISION, ALLOCATABLE :: dv(:)
      IF (ALLOCATED(coord%dir)) THEN
         PRINT*, "[WARNING] Multiple allocations ", &
             "attempted at load_dir_tables. Skipping..."
         RETURN
      ENDIF
      ALLOCATE(coord%dir(coord%C,coord%d), dv(coord%C*coord%d))
      CALL rd_check(rd_load_dirichlet_tables(TRIM(tdir)//C_NULL_CHAR, &
           INT(coord%C,C_INT),INT(coord%d,C_INT),dv),'load_dirichlet_tables')
      coord%dir = RESHAPE(dv,(/coord%C,coord%d/))
      DEALLOCATE(dv)
      END SUBROUTINE load_dirichlet_tables

! The Neumann reconstructor of order ord (1: first, 2: second normal
! derivative) of coord, from Q{d}.dat and Q{ord}n{d}.dat
! (fcgram_mod.f90:267-371).
      SUBROUTINE load_neumann_tables(coord,tdir,ord)
      TYPE(FCCOORD), INTENT(INOUT) :: coord
      CHARACTER(len=*), INTENT(IN) :: tdir
      INTEGER, INTENT(IN)          :: ord
      DOUBLE PRECISION, ALLOCATABLE :: w(:)
      IF (ord .eq. 1) THEN
         IF (ALLOCATED(coord%neu)) THEN
            PRINT*, "[WARNING] Multiple allocations ", &
                "attempted at load_neu_tables. Skipping..."
            RETURN
         ENDIF
      ELSEIF (ord .eq. 2) THEN
         IF (ALLOCATED(coord%neu2)) THEN
            PRINT*, "[WARNING] Multiple allocations ", &
                "attempted at load_neu_tables. Skipping..."
            RETURN
         ENDIF
      ELSE
         PRINT*, "[WARNING] load_neu_tables called with ", &
             "undefined order. Skipping..."
         RETURN
      ENDIF
      ALLOCATE(w(coord%d))
      CALL rd_check(rd_load_neumann_tables(TRIM(tdir)//C_NULL_CHAR,INT(coord%d,C_INT), &
           INT(ord,C_INT),REAL(coord%dx,C_DOUBLE),w),'load_neumann_tables')
      IF (ord .eq. 1) THEN
         CALL MOVE_ALLOC(w,coord%neu)
      ELSE
         CALL MOVE_ALLOC(w,coord%neu2)
      ENDIF
      END SUBROUTINE load_neumann_tables

! Wall value from its ord-th outward normal derivative, boundary 5
! (z(1)) or 6 (z(nz-Cz)) (fcgram_mod.f90:374-517).  f is a spectral
! (nlm,nz) array with a device copy: z is its second index here, the
! radial one, and the boundary row holds the derivative on entry.
      SUBROUTINE neumann_reconstruct(plan,f,boun,ord)
      TYPE(FCPLAN), INTENT(IN)              :: plan
      DOUBLE COMPLEX, INTENT(INOUT), TARGET :: f(*)
      INTEGER, INTENT(IN)                   :: boun, ord
      LOGICAL :: ok
      ok = (boun .eq. 5 .OR. boun .eq. 6) .AND. &
           ((ord .eq. 1 .AND. ALLOCATED(plan%z%neu)) .OR. &
            (ord .eq. 2 .AND. ALLOCATED(plan%z%neu2)))
      IF (.NOT. ok) THEN
         PRINT*, "[ERROR] Neumann reconstruction not ", &
             "performed. Unsupported derivative order, boundary surface or ", &
             "unallocated table. Aborting..."
         STOP
      ENDIF
      IF (ord .eq. 1) THEN
         CALL rd_check(rd_neumann_reconstruct(plan%h,rd_dptr(C_LOC(f),'f'),INT(boun,C_INT), &
              1_C_INT,plan%z%neu),'neumann_reconstruct')
      ELSE
         CALL rd_check(rd_neumann_reconstruct(plan%h,rd_dptr(C_LOC(f),'f'),INT(boun,C_INT), &
              2_C_INT,plan%z%neu2),'neumann_reconstruct')
      ENDIF
      CALL fc_sync(plan)
      END SUBROUTINE neumann_reconstruct

! Wall value from the Robin condition f' + a f = g (fcgram_mod.f90:520-650);
! a holds one real coefficient per lm (nlm values in any shape).
      SUBROUTINE robin_reconstruct(plan,f,boun,a)
      TYPE(FCPLAN), INTENT(IN)                :: plan
      DOUBLE COMPLEX, INTENT(INOUT), TARGET   :: f(*)
      DOUBLE PRECISION, INTENT(INOUT), TARGET :: a(:,:)
      INTEGER, INTENT(IN)                     :: boun
      INTEGER(C_SIZE_T) :: nb
      LOGICAL :: own
      IF (boun .eq. 5 .OR. boun .eq. 6) THEN
         IF (.NOT. ALLOCATED(plan%z%neu)) THEN
            PRINT*, "[ERROR] Neumann table not loaded in", &
                "robin_reconstruct. Aborting..."
            STOP
         ENDIF
      ELSE
         PRINT*, "[ERROR] Robin reconstruction not ", &
             "performed. Wrong boundary specified. Aborting..."
         STOP
      ENDIF
      IF (SIZE(a) .ne. plan%planrc%nlm) THEN
         PRINT*, "[ERROR] robin_reconstruct: one coefficient per lm expected"
         STOP
      ENDIF
      nb = 8_C_SIZE_T*SIZE(a,KIND=C_SIZE_T)
      own = rd_dev_present(C_LOC(a)) .eq. 0
      IF (own) CALL rd_check(rd_dev_alloc(C_LOC(a),nb),'robin_reconstruct')
      CALL rd_check(rd_dev_update_to(C_LOC(a),nb),'robin_reconstruct')
      CALL rd_check(rd_robin_reconstruct(plan%h,rd_dptr(C_LOC(f),'f'),INT(boun,C_INT), &
           plan%z%neu,rd_dev_lookup(C_LOC(a))),'robin_reconstruct')
      CALL fc_sync(plan)
      IF (own) CALL rd_check(rd_dev_free(C_LOC(a)),'robin_reconstruct')
      END SUBROUTINE robin_reconstruct

! Waits for the plan's queued device work (each reference routine is
! synchronous).
      SUBROUTINE fc_sync(this)
      TYPE(FCPLAN), INTENT(IN) :: this
      CALL rd_check(rd_plan_sync(this%h),'device_sync')
      END SUBROUTINE fc_sync

      END MODULE fcgram

!=================================================================
      MODULE fft
      USE fcgram
      TYPE(FCPLAN), SAVE, TARGET :: planfc
      END MODULE fft

!=================================================================
! grid, kes, boxsize -- the grid modules of src/pseudo/pseudospec_mod.f90
! (:13-69, :111-117, :119-236) without SHTns: box_init takes the Gauss
! nodes and the (l,m) of each lm from a library plan made only for that,
! with the grid of the path's plan, as the reference does with a
! throw-away SHTns configuration.  One rank holds every plane.
      MODULE grid
      IMPLICIT NONE
      INTEGER :: lmax
      INTEGER :: nlm
      INTEGER :: nlat
      INTEGER :: nphi
      INTEGER :: nz
      INTEGER :: Cz = 0
      INTEGER :: oz = 0
      DOUBLE PRECISION, ALLOCATABLE, DIMENSION(:) :: theta,phi,z
      SAVE
      CONTAINS

! Reads lmax, nz, Cz, oz from the namelist 'grid' of infile_ and applies the
! 3/2 rule: nlat is the smallest multiple of 4 not below 3(lmax+1)/2, nphi =
! 2 nlat (pseudospec_mod.f90:41-69).
      SUBROUTINE grid_init(infile_)
      CHARACTER(len=*), INTENT(IN) :: infile_
      NAMELIST / grid / lmax,nz,Cz,oz
      Cz = 0
      oz = 0
      OPEN(1,file=infile_,status='unknown',form="formatted")
      READ(1,NML=grid)
      CLOSE(1)
      nlm  = (lmax+1)*(lmax+2)/2
      nlat = 4*((3*(lmax+1)+7)/8)
      nphi = 2*nlat
      END SUBROUTINE grid_init
      END MODULE grid

      MODULE kes
      IMPLICIT NONE
      DOUBLE PRECISION, ALLOCATABLE, DIMENSION(:) :: kz
      INTEGER, ALLOCATABLE, DIMENSION(:)          :: ell,emm
      SAVE
      END MODULE kes

      MODULE boxsize
      IMPLICIT NONE
      DOUBLE PRECISION :: rint,rext
      DOUBLE PRECISION :: Dkz
      SAVE
      CONTAINS

! Reads rint, rext from the namelist 'boxparams' of infile_ and builds the
! grids: theta at the Gauss nodes (north to south), phi, the uniform z from
! rint (z(1)) to rext (z(nz-Cz)) followed by the Cz continuation points,
! kz in FFT order (for odd nz the symmetric order 0..(nz-1)/2,
! -(nz-1)/2..-1, which the reference's loop leaves undefined), ell and emm;
! writes theta.txt, phi.txt and z.txt (pseudospec_mod.f90:126-236).
      SUBROUTINE box_init(infile_)
      USE iso_c_binding
      USE rd_capi
      USE grid
      USE kes
      CHARACTER(len=*), INTENT(IN) :: infile_
      NAMELIST / boxparams / rint,rext
      DOUBLE PRECISION, PARAMETER :: pi = 3.14159265358979323846D0
      DOUBLE PRECISION, ALLOCATABLE, TARGET :: ct(:)
      INTEGER(C_INT), ALLOCATABLE, TARGET   :: li(:),mi(:)
      TYPE(C_PTR) :: h
      DOUBLE PRECISION :: rmp
      INTEGER :: i,j,k
      OPEN(1,file=infile_,status='unknown',form="formatted")
      READ(1,NML=boxparams)
      CLOSE(1)
      IF (ALLOCATED(theta)) DEALLOCATE(theta,phi,z,kz,ell,emm)
      ALLOCATE(theta(nlat),phi(nphi),z(nz),kz(nz),ell(nlm),emm(nlm))
      ALLOCATE(ct(nlat),li(nlm),mi(nlm))
      h = C_NULL_PTR
      CALL rd_check(rd_plan_create(h,INT(lmax,C_INT),INT(nlat,C_INT),INT(nphi,C_INT), &
           4_C_INT,0_C_INT,0_C_INT,1.0D0,0.0D0,C_NULL_PTR),'box_init')
      CALL rd_check(rd_plan_grid(h,C_LOC(ct),C_NULL_PTR,C_NULL_PTR,C_NULL_PTR,C_NULL_PTR, &
           C_LOC(li),C_LOC(mi)),'box_init')
      CALL rd_check(rd_plan_destroy(h),'box_init')
      DO i = 1,nlat
         theta(i) = ACOS(ct(i))
      END DO
      DO i = 1,nlm
         ell(i) = li(i)
         emm(i) = mi(i)
      END DO
      DEALLOCATE(ct,li,mi)
      rmp = 2.0D0*pi/nphi
      DO j = 1,nphi
         phi(j) = rmp*(j-1)
      ENDDO
      rmp = (rext-rint)/(nz-Cz-1)
      Dkz = 2.0D0*pi/(rmp*nz)
      DO k = 1,nz
         z(k) = rint+rmp*(k-1)
      ENDDO
      OPEN(1,file='theta.txt',action='write')
      OPEN(2,file='phi.txt',action='write')
      OPEN(3,file='z.txt',action='write')
      DO i = 1,nlat
         WRITE(1,FMT='(1P E23.15)') theta(i)
      ENDDO
      DO i = 1,nphi
         WRITE(2,FMT='(1P E23.15)') phi(i)
      ENDDO
      DO i = 1,nz-Cz
         WRITE(3,FMT='(1P E23.15)') z(i)
      ENDDO
      CLOSE(1)
      CLOSE(2)
      CLOSE(3)
      IF (MOD(nz,2) .eq. 0) THEN
         DO k = 1,nz/2
            kz(k) = REAL(k-1,8)
            kz(k+nz/2) = REAL(k-nz/2-1,8)
         END DO
      ELSE
         DO k = 1,nz
            IF (k-1 .le. (nz-1)/2) THEN
               kz(k) = REAL(k-1,8)
            ELSE
               kz(k) = REAL(k-1-nz,8)
            ENDIF
         END DO
      ENDIF
      kz = kz*Dkz
      END SUBROUTINE box_init
      END MODULE boxsize

!=================================================================
! boundary -- TYPE(BCPLAN), setup_bc, v_parsebc, v_setup, dirichlet,
! derivr_mat, invDl_setup, invDl, v_imposebc, v_imposebc_fast
! (src/boundary/boundary_mod.f90, src/boundary/vboundary.f90).  The
! tables (binv, d2bor) live in the plan on the device; the routines use
! the global plan planfc, as the reference's do.
      MODULE boundary
      USE iso_c_binding
      USE rd_capi
      USE fcgram
      IMPLICIT NONE

      TYPE BCPLAN
         INTEGER :: bcint, bcext
      END TYPE BCPLAN

      CONTAINS

      SUBROUTINE setup_bc(this,planfc,bckind,field)
      TYPE(BCPLAN), INTENT(INOUT)   :: this
      TYPE(FCPLAN), INTENT(INOUT)   :: planfc
      CHARACTER(len=15), INTENT(IN) :: bckind(2)
      CHARACTER(len=*), INTENT(IN)  :: field
      IF (field .eq. 'v') CALL v_setup(this,planfc,bckind)
      END SUBROUTINE setup_bc

      FUNCTION v_parsebc(str) RESULT(bc)
      CHARACTER(len=*), INTENT(IN) :: str
      INTEGER :: bc
      IF (str .eq. 'noslip') THEN
         bc = 0
      ELSE
         PRINT*, "[ERROR] Unknown boundary ", "condition type", str, " Aborting..."
         STOP
      ENDIF
      END FUNCTION v_parsebc

      SUBROUTINE v_setup(this,planfc,bckind)
      TYPE(BCPLAN), INTENT(INOUT)   :: this
      TYPE(FCPLAN), INTENT(INOUT)   :: planfc
      CHARACTER(len=15), INTENT(IN) :: bckind(2)
      this%bcint = v_parsebc(lowtrim(bckind(1)))
      this%bcext = v_parsebc(lowtrim(bckind(2)))
      CALL rd_check(rd_bc_setup(planfc%h,INT(this%bcint,C_INT),INT(this%bcext,C_INT)), &
           'invDl_setup')
      END SUBROUTINE v_setup

! The M_l of both walls without slip, on the global plan.
      SUBROUTINE invDl_setup()
      USE fft
      CALL rd_check(rd_bc_setup(planfc%h,0_C_INT,0_C_INT),'invDl_setup')
      END SUBROUTINE invDl_setup

      SUBROUTINE dirichlet(a)
      USE fft
      DOUBLE COMPLEX, INTENT(INOUT), TARGET :: a(*)
      CALL rd_check(rd_dirichlet(planfc%h,rd_dptr(C_LOC(a),'a')),'dirichlet')
      CALL fc_sync(planfc)
      END SUBROUTINE dirichlet

! d(i,j) = d^p e_j / dr^p at z(i) on the N = nz - Cz physical points;
! d is a host array, filled through a temporary device copy.
      SUBROUTINE derivr_mat(p,d)
      USE fft
      INTEGER, INTENT(IN) :: p
      DOUBLE PRECISION, INTENT(OUT), TARGET :: d(:,:)
      INTEGER(C_SIZE_T) :: nb
      nb = 8_C_SIZE_T*SIZE(d,KIND=C_SIZE_T)
      CALL rd_check(rd_dev_alloc(C_LOC(d),nb),'derivr_mat')
      CALL rd_check(rd_derivr_mat(planfc%h,INT(p,C_INT),rd_dev_lookup(C_LOC(d))),'derivr_mat')
      CALL rd_check(rd_dev_update_from(C_LOC(d),nb),'derivr_mat')
      CALL rd_check(rd_dev_free(C_LOC(d)),'derivr_mat')
      END SUBROUTINE derivr_mat

      SUBROUTINE invDl(a,b)
      USE fft
      DOUBLE COMPLEX, INTENT(IN), TARGET  :: a(*)
      DOUBLE COMPLEX, INTENT(OUT), TARGET :: b(*)
      CALL rd_check(rd_invdl(planfc%h,rd_dptr(C_LOC(a),'a'),rd_dptr(C_LOC(b),'b')),'invDl')
      CALL fc_sync(planfc)
      END SUBROUTINE invDl

      SUBROUTINE v_imposebc(tor,phi,pol)
      USE fft
      DOUBLE COMPLEX, INTENT(INOUT), TARGET :: tor(*),phi(*)
      DOUBLE COMPLEX, INTENT(OUT), TARGET   :: pol(*)
      CALL rd_check(rd_v_imposebc(planfc%h,rd_dptr(C_LOC(tor),'tor'), &
           rd_dptr(C_LOC(phi),'phi'),rd_dptr(C_LOC(pol),'pol')),'v_imposebc')
      CALL fc_sync(planfc)
      END SUBROUTINE v_imposebc

      SUBROUTINE v_imposebc_fast(tor,phi,pol)
      USE fft
      DOUBLE COMPLEX, INTENT(INOUT), TARGET :: tor(*),phi(*)
      DOUBLE COMPLEX, INTENT(OUT), TARGET   :: pol(*)
      CALL rd_check(rd_v_imposebc_fast(planfc%h,rd_dptr(C_LOC(tor),'tor'), &
           rd_dptr(C_LOC(phi),'phi'),rd_dptr(C_LOC(pol),'pol')),'v_imposebc_fast')
      CALL fc_sync(planfc)
      END SUBROUTINE v_imposebc_fast

! Lower case, without leading or trailing blanks (the reference's
! preprocess/lowcase).
      FUNCTION lowtrim(s) RESULT(o)
      CHARACTER(len=*), INTENT(IN) :: s
      CHARACTER(len=LEN(s)) :: o
      INTEGER :: j
      o = TRIM(ADJUSTL(s))
      DO j = 1,LEN(o)
         IF (o(j:j) >= 'A' .AND. o(j:j) <= 'Z') o(j:j) = ACHAR(IACHAR(o(j:j))+32)
      END DO
      END FUNCTION lowtrim

      END MODULE boundary
