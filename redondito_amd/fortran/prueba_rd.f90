!=================================================================
! Known-answer checks of the Fortran drop-in layer, in the style of the
! reference's prueba_sphp / prueba_vec / prueba_z / prueba_3d programs:
! host arrays with registered device copies, the reference's routine
! names and argument lists, exact answers from closed forms.  Prints one
! CHECK line per case and stops with code 1 if any case fails.
!=================================================================
      PROGRAM prueba_rd
      USE iso_c_binding
      USE gdevice
      USE fcgram
      USE fft
      IMPLICIT NONE
      DOUBLE PRECISION, PARAMETER :: pi = 3.14159265358979323846D0
      INTEGER :: nfail = 0

      CALL device_init(0)
      CALL casos_sht()
      CALL casos_vec()
      CALL caso_z()
      CALL caso_uxw()
      CALL caso_coupled()
      CALL caso_boundary()
      CALL caso_box()
      CALL caso_neumann()
      CALL caso_redist()
      IF (nfail .gt. 0) THEN
         PRINT '(A,I0,A)', 'prueba_rd: ', nfail, ' failed'
         STOP 1
      ENDIF
      PRINT '(A)', 'prueba_rd: all passed'

      CONTAINS

      SUBROUTINE report(name,err,tol)
      CHARACTER(len=*), INTENT(IN) :: name
      DOUBLE PRECISION, INTENT(IN) :: err,tol
      IF (err .le. tol) THEN
         PRINT '(A,A40,ES10.2,A,ES8.1,A)', 'CHECK ', name, err, ' <= ', tol, ' PASS'
      ELSE
         PRINT '(A,A40,ES10.2,A,ES8.1,A)', 'CHECK ', name, err, ' >  ', tol, ' FAIL'
         nfail = nfail + 1
      ENDIF
      END SUBROUTINE report

      INTEGER FUNCTION lm(lmax,l,m)
      INTEGER, INTENT(IN) :: lmax,l,m
      lm = m*(2*lmax+1-m)/2 + l + 1
      END FUNCTION lm

! The grid of a plan: cos and sin of the Gauss nodes.
      SUBROUTINE grilla(plan,ct,st)
      TYPE(FCPLAN), INTENT(IN) :: plan
      DOUBLE PRECISION, TARGET, INTENT(OUT) :: ct(:),st(:)
      CALL rd_check(rd_plan_grid(plan%h,C_LOC(ct),C_LOC(st),C_NULL_PTR,C_NULL_PTR, &
           C_NULL_PTR,C_NULL_PTR,C_NULL_PTR),'grid')
      END SUBROUTINE grilla

! prueba_sphp cases 3, 4, 8 and 9 (src/prueba_sphp.f90:121-184), plane k x k
      SUBROUTINE casos_sht()
      INTEGER, PARAMETER :: lmax = 15, nz = 6, nlat = 24, nphi = 48
      INTEGER, PARAMETER :: nlm = (lmax+1)*(lmax+2)/2
      TYPE(FCPLAN) :: plan
      DOUBLE PRECISION, TARGET :: f3(nlat,nphi,nz), ct(nlat), st(nlat)
      DOUBLE COMPLEX, TARGET :: b3(nlm,nz)
      DOUBLE COMPLEX :: a(nlm)
      DOUBLE PRECISION :: f(nlat,nphi), c, ph, err1, err2
      INTEGER :: ic, i, j, k
      CHARACTER(len=40) :: nombre
      CALL fcgram_create_plan(plan,(/nlat,nphi,nz/),(/0,0,0/),(/0,0,0/), &
           (/1.0D0,1.0D0,1.0D0/),'.',0,lmax)
      CALL grilla(plan,ct,st)
      CALL gdev_alloc(C_LOC(f3),INT(8*SIZE(f3),C_SIZE_T))
      CALL gdev_alloc(C_LOC(b3),INT(16*SIZE(b3),C_SIZE_T))
      DO ic = 1,4
         a = (0.0D0,0.0D0)
         DO j = 1,nphi
            ph = 2*pi*(j-1)/nphi
            DO i = 1,nlat
               SELECT CASE (ic)
               CASE (1)
                  f(i,j) = st(i)*COS(ph)
               CASE (2)
                  f(i,j) = st(i)*SIN(ph)
               CASE (3)
                  f(i,j) = st(i)**lmax*COS(lmax*ph)
               CASE (4)
                  f(i,j) = 2-3*ct(i)+st(i)*SIN(ph)+0.5D0*st(i)**2*COS(2*ph)
               END SELECT
            END DO
         END DO
         SELECT CASE (ic)
         CASE (1)
            nombre = 'sht sin(th) cos(ph)'
            a(lm(lmax,1,1)) = SQRT(2*pi/3)
         CASE (2)
            nombre = 'sht sin(th) sin(ph)'
            a(lm(lmax,1,1)) = DCMPLX(0.0D0,-SQRT(2*pi/3))
         CASE (3)
            nombre = 'sht sin(th)^lmax cos(lmax ph)'
            c = SQRT((2*lmax+1)/(4*pi))
            DO k = 1,lmax
               c = c*SQRT((2*k-1.0D0)/(2*k))
            END DO
            a(lm(lmax,lmax,lmax)) = 1/(2*c)
         CASE (4)
            nombre = 'sht sum 2(1) - 3(2) + (4) + (6)/2'
            a(lm(lmax,0,0)) = 2*SQRT(4*pi)
            a(lm(lmax,1,0)) = -3*SQRT(4*pi/3)
            a(lm(lmax,1,1)) = DCMPLX(0.0D0,-SQRT(2*pi/3))
            a(lm(lmax,2,2)) = SQRT(2*pi/15)
         END SELECT
         DO k = 1,nz
            f3(:,:,k) = k*f
         END DO
         CALL gdev_update_to(C_LOC(f3),INT(8*SIZE(f3),C_SIZE_T))
         CALL fftp2d_real_to_complex_xy(plan,f3,b3,0)
         CALL gdev_update_from(C_LOC(b3),INT(16*SIZE(b3),C_SIZE_T))
         err1 = 0
         DO k = 1,nz
            err1 = MAX(err1,MAXVAL(ABS(b3(:,k)-k*a)))
         END DO
         DO k = 1,nz
            b3(:,k) = k*a
         END DO
         CALL gdev_update_to(C_LOC(b3),INT(16*SIZE(b3),C_SIZE_T))
         CALL fftp2d_complex_to_real_xy(plan,b3,f3,0)
         CALL gdev_update_from(C_LOC(f3),INT(8*SIZE(f3),C_SIZE_T))
         err2 = 0
         DO k = 1,nz
            err2 = MAX(err2,MAXVAL(ABS(f3(:,:,k)-k*f)))
         END DO
         CALL report(TRIM(nombre)//' analysis',err1,1.0D-12)
         CALL report(TRIM(nombre)//' synthesis',err2,1.0D-12)
      END DO
      CALL gdev_free(C_LOC(f3))
      CALL gdev_free(C_LOC(b3))
      CALL fcgram_destroy_plan(plan)
      END SUBROUTINE casos_sht

! prueba_vec case 8 (src/prueba_vec.f90:198-207): Q = cos th, S = T = sin th cos ph
      SUBROUTINE casos_vec()
      INTEGER, PARAMETER :: lmax = 15, nz = 6, nlat = 24, nphi = 48
      INTEGER, PARAMETER :: nlm = (lmax+1)*(lmax+2)/2
      TYPE(FCPLAN) :: plan
      DOUBLE PRECISION, TARGET :: vr(nlat,nphi,nz), vt(nlat,nphi,nz), vp(nlat,nphi,nz)
      DOUBLE PRECISION, TARGET :: ct(nlat), st(nlat)
      DOUBLE COMPLEX, TARGET :: q(nlm,nz), s(nlm,nz), t(nlm,nz)
      DOUBLE COMPLEX :: aq(nlm), as(nlm), at(nlm)
      DOUBLE PRECISION :: ph, err1, err2
      INTEGER :: i, j, k
      CALL fcgram_create_plan(plan,(/nlat,nphi,nz/),(/0,0,0/),(/0,0,0/), &
           (/1.0D0,1.0D0,1.0D0/),'.',0,lmax)
      CALL grilla(plan,ct,st)
      CALL gdev_alloc(C_LOC(vr),INT(8*SIZE(vr),C_SIZE_T))
      CALL gdev_alloc(C_LOC(vt),INT(8*SIZE(vt),C_SIZE_T))
      CALL gdev_alloc(C_LOC(vp),INT(8*SIZE(vp),C_SIZE_T))
      CALL gdev_alloc(C_LOC(q),INT(16*SIZE(q),C_SIZE_T))
      CALL gdev_alloc(C_LOC(s),INT(16*SIZE(s),C_SIZE_T))
      CALL gdev_alloc(C_LOC(t),INT(16*SIZE(t),C_SIZE_T))
      aq = 0; as = 0; at = 0
      aq(lm(lmax,1,0)) = SQRT(4*pi/3)
      as(lm(lmax,1,1)) = SQRT(2*pi/3)
      at(lm(lmax,1,1)) = SQRT(2*pi/3)
      DO k = 1,nz
         DO j = 1,nphi
            ph = 2*pi*(j-1)/nphi
            DO i = 1,nlat
               vr(i,j,k) = k*ct(i)
               vt(i,j,k) = k*(ct(i)*COS(ph)-SIN(ph))
               vp(i,j,k) = k*(-SIN(ph)-ct(i)*COS(ph))
            END DO
         END DO
      END DO
      CALL gdev_update_to(C_LOC(vr),INT(8*SIZE(vr),C_SIZE_T))
      CALL gdev_update_to(C_LOC(vt),INT(8*SIZE(vt),C_SIZE_T))
      CALL gdev_update_to(C_LOC(vp),INT(8*SIZE(vp),C_SIZE_T))
      CALL fftp2d_vec_to_qst(plan,vr,vt,vp,q,s,t,0)
      CALL gdev_update_from(C_LOC(q),INT(16*SIZE(q),C_SIZE_T))
      CALL gdev_update_from(C_LOC(s),INT(16*SIZE(s),C_SIZE_T))
      CALL gdev_update_from(C_LOC(t),INT(16*SIZE(t),C_SIZE_T))
      err1 = 0
      DO k = 1,nz
         err1 = MAX(err1,MAXVAL(ABS(q(:,k)-k*aq)),MAXVAL(ABS(s(:,k)-k*as)),MAXVAL(ABS(t(:,k)-k*at)))
      END DO
      CALL fftp2d_qst_to_vec(plan,q,s,t,vr,vt,vp,0)
      CALL gdev_update_from(C_LOC(vt),INT(8*SIZE(vt),C_SIZE_T))
      err2 = 0
      DO k = 1,nz
         DO j = 1,nphi
            ph = 2*pi*(j-1)/nphi
            err2 = MAX(err2,MAXVAL(ABS(vt(:,j,k)-k*(ct*COS(ph)-SIN(ph)))))
         END DO
      END DO
      CALL report('vec sum of cases 1, 4, 5 analysis',err1,1.0D-12)
      CALL report('vec round trip v_theta',err2,1.0D-12)
      CALL gdev_free(C_LOC(vr)); CALL gdev_free(C_LOC(vt)); CALL gdev_free(C_LOC(vp))
      CALL gdev_free(C_LOC(q)); CALL gdev_free(C_LOC(s)); CALL gdev_free(C_LOC(t))
      CALL fcgram_destroy_plan(plan)
      END SUBROUTINE casos_vec

! prueba_z (b) at nz = 128 (src/prueba_z.f90:79-125) through derivr of
! the global plan: published error 4.9e-8 (LEEME.md:34).
      SUBROUTINE caso_z()
      INTEGER, PARAMETER :: lmax = 7, nz = 128, C = 25, od = 5
      INTEGER, PARAMETER :: nlm = (lmax+1)*(lmax+2)/2, nlat = 12, nphi = 24
      DOUBLE COMPLEX, TARGET :: f(nlm,nz), e(nlm,nz), d(nlm,nz)
      DOUBLE PRECISION :: dz, zj, err
      INTEGER :: j
      CHARACTER(len=256) :: tdir
      CALL GET_ENVIRONMENT_VARIABLE('RD_TABLE_DIR',tdir)
      dz = 1.0D0/(nz-C-1)
      CALL fcgram_create_plan(planfc,(/nlat,nphi,nz/),(/0,0,C/),(/0,0,od/), &
           (/1.0D0,1.0D0,dz/),TRIM(tdir),0,lmax)
      f = 0; e = 0
      DO j = 1,nz-C
         zj = (j-1)*dz
         f(:,j) = DCMPLX(0.3D0,-0.5D0)*(SIN(2.3D0*zj+0.4D0)+zj**2)
         e(:,j) = DCMPLX(0.3D0,-0.5D0)*(2.3D0*COS(2.3D0*zj+0.4D0)+2*zj)
      END DO
      CALL gdev_alloc(C_LOC(f),INT(16*SIZE(f),C_SIZE_T))
      CALL gdev_alloc(C_LOC(d),INT(16*SIZE(d),C_SIZE_T))
      CALL gdev_update_to(C_LOC(f),INT(16*SIZE(f),C_SIZE_T))
      CALL derivr(f,d,1,'r')
      CALL gdev_update_from(C_LOC(d),INT(16*SIZE(d),C_SIZE_T))
      err = MAXVAL(ABS(d(:,1:nz-C)-e(:,1:nz-C)))/MAXVAL(ABS(e(:,1:nz-C)))
      CALL report('derivr FC-Gram nz=128 (published 4.9e-8)',err,1.0D-7)
      CALL gdev_free(C_LOC(f)); CALL gdev_free(C_LOC(d))
      CALL fcgram_destroy_plan(planfc)
      END SUBROUTINE caso_z

! u x omega of a solid-body rotation: NQ_00, NQ_20, NS_20 in closed form.
      SUBROUTINE caso_uxw()
      INTEGER, PARAMETER :: lmax = 15, nr = 64, C = 25, od = 5, nz = nr + C
      INTEGER, PARAMETER :: nlat = 24, nphi = 48, nlm = (lmax+1)*(lmax+2)/2
      DOUBLE PRECISION, PARAMETER :: ri = 0.35D0, ro = 1.0D0, om = 0.7D0
      DOUBLE COMPLEX, TARGET :: w(nlm,nz), z(nlm,nz), nq(nlm,nz), ns(nlm,nz), nt(nlm,nz)
      DOUBLE PRECISION :: dz, r, eq, es, et, sc
      INTEGER :: k
      CHARACTER(len=256) :: tdir
      CALL GET_ENVIRONMENT_VARIABLE('RD_TABLE_DIR',tdir)
      dz = (ro-ri)/(nr-1)
      CALL fcgram_create_plan(planfc,(/nlat,nphi,nz/),(/0,0,C/),(/0,0,od/), &
           (/1.0D0,1.0D0,dz/),TRIM(tdir),0,lmax,ri)
      w = 0; z = 0
      DO k = 1,nr
         r = ri+(k-1)*dz
         z(lm(lmax,1,0),k) = om*r**2*SQRT(4*pi/3)
      END DO
      CALL gdev_alloc(C_LOC(w),INT(16*SIZE(w),C_SIZE_T))
      CALL gdev_alloc(C_LOC(z),INT(16*SIZE(z),C_SIZE_T))
      CALL gdev_alloc(C_LOC(nq),INT(16*SIZE(nq),C_SIZE_T))
      CALL gdev_alloc(C_LOC(ns),INT(16*SIZE(ns),C_SIZE_T))
      CALL gdev_alloc(C_LOC(nt),INT(16*SIZE(nt),C_SIZE_T))
      CALL gdev_update_to(C_LOC(w),INT(16*SIZE(w),C_SIZE_T))
      CALL gdev_update_to(C_LOC(z),INT(16*SIZE(z),C_SIZE_T))
      CALL nonlinear_uxw(w,z,nq,ns,nt,0)
      CALL gdev_update_from(C_LOC(nq),INT(16*SIZE(nq),C_SIZE_T))
      CALL gdev_update_from(C_LOC(ns),INT(16*SIZE(ns),C_SIZE_T))
      CALL gdev_update_from(C_LOC(nt),INT(16*SIZE(nt),C_SIZE_T))
      eq = 0; es = 0; et = 0; sc = 0
      DO k = 1,nr
         r = ri+(k-1)*dz
         sc = MAX(sc,2*om**2*r*(2.0D0/3)*SQRT(4*pi))
         eq = MAX(eq,ABS(nq(lm(lmax,0,0),k)-2*om**2*r*(2.0D0/3)*SQRT(4*pi)), &
                  ABS(nq(lm(lmax,2,0),k)+2*om**2*r*(2.0D0/3)*SQRT(4*pi/5)))
         es = MAX(es,ABS(ns(lm(lmax,2,0),k)+om**2*r*(2.0D0/3)*SQRT(4*pi/5)))
         et = MAX(et,MAXVAL(ABS(nt(:,k))))
      END DO
      CALL report('u x omega solid body NQ (rel)',eq/sc,1.0D-9)
      CALL report('u x omega solid body NS (rel)',es/sc,1.0D-12)
      CALL report('u x omega solid body NT (abs)',et,1.0D-12)
      CALL gdev_free(C_LOC(w)); CALL gdev_free(C_LOC(z))
      CALL gdev_free(C_LOC(nq)); CALL gdev_free(C_LOC(ns)); CALL gdev_free(C_LOC(nt))
      CALL fcgram_destroy_plan(planfc)
      END SUBROUTINE caso_uxw

! The coupled step for a solid-body rotation in a uniform field B0 z:
! u x B = grad(Omega B0 r^2 sin^2 / 2) (MQ_00, MQ_20, MS_20) and
! u . grad T = Omega dT/dphi (NH_lm = i m Omega T_lm) for a deterministic T.
      SUBROUTINE caso_coupled()
      INTEGER, PARAMETER :: lmax = 15, nr = 64, C = 25, od = 5, nz = nr + C
      INTEGER, PARAMETER :: nlat = 24, nphi = 48, nlm = (lmax+1)*(lmax+2)/2
      DOUBLE PRECISION, PARAMETER :: ri = 0.35D0, ro = 1.0D0, om = 0.7D0, b0 = 1.3D0
      DOUBLE COMPLEX, TARGET :: w(nlm,nz), z(nlm,nz), pb(nlm,nz), tb(nlm,nz), t(nlm,nz)
      DOUBLE COMPLEX, TARGET :: nq(nlm,nz), ns(nlm,nz), nt(nlm,nz), mq(nlm,nz), ms(nlm,nz)
      DOUBLE COMPLEX, TARGET :: mt(nlm,nz), nh(nlm,nz)
      DOUBLE PRECISION :: dz, r, eq, es, et, eh, sc, sh
      INTEGER :: k, l, m
      CHARACTER(len=256) :: tdir
      CALL GET_ENVIRONMENT_VARIABLE('RD_TABLE_DIR',tdir)
      dz = (ro-ri)/(nr-1)
      CALL fcgram_create_plan(planfc,(/nlat,nphi,nz/),(/0,0,C/),(/0,0,od/), &
           (/1.0D0,1.0D0,dz/),TRIM(tdir),0,lmax,ri)
      w = 0; z = 0; pb = 0; tb = 0; t = 0
      DO k = 1,nr
         r = ri+(k-1)*dz
         z(lm(lmax,1,0),k) = om*r**2*SQRT(4*pi/3)
         pb(lm(lmax,1,0),k) = 0.5D0*b0*r**2*SQRT(4*pi/3)
         DO m = 0,lmax
            DO l = m,lmax
               t(lm(lmax,l,m),k) = SIN(pi*(r-ri)/(ro-ri))/(l+1.0D0)**1.5D0 * &
                    DCMPLX(COS(1.0D0*(l+m)),MERGE(0.0D0,SIN(1.0D0*l*m),m.eq.0))
            END DO
         END DO
      END DO
      CALL dev_all(w); CALL dev_all(z); CALL dev_all(pb); CALL dev_all(tb); CALL dev_all(t)
      CALL dev_all(nq); CALL dev_all(ns); CALL dev_all(nt); CALL dev_all(mq); CALL dev_all(ms)
      CALL dev_all(mt); CALL dev_all(nh)
      CALL gdev_update_to(C_LOC(w),INT(16*SIZE(w),C_SIZE_T))
      CALL gdev_update_to(C_LOC(z),INT(16*SIZE(z),C_SIZE_T))
      CALL gdev_update_to(C_LOC(pb),INT(16*SIZE(pb),C_SIZE_T))
      CALL gdev_update_to(C_LOC(tb),INT(16*SIZE(tb),C_SIZE_T))
      CALL gdev_update_to(C_LOC(t),INT(16*SIZE(t),C_SIZE_T))
      CALL nonlinear_coupled(w,z,pb,tb,t,nq,ns,nt,mq,ms,mt,nh,0)
      CALL gdev_update_from(C_LOC(mq),INT(16*SIZE(mq),C_SIZE_T))
      CALL gdev_update_from(C_LOC(ms),INT(16*SIZE(ms),C_SIZE_T))
      CALL gdev_update_from(C_LOC(mt),INT(16*SIZE(mt),C_SIZE_T))
      CALL gdev_update_from(C_LOC(nh),INT(16*SIZE(nh),C_SIZE_T))
      eq = 0; es = 0; et = 0; eh = 0; sc = 0; sh = 0
      DO k = 1,nr
         r = ri+(k-1)*dz
         sc = MAX(sc,(2.0D0/3)*om*b0*r*SQRT(4*pi))
         eq = MAX(eq,ABS(mq(lm(lmax,0,0),k)-(2.0D0/3)*om*b0*r*SQRT(4*pi)), &
                  ABS(mq(lm(lmax,2,0),k)+(2.0D0/3)*om*b0*r*SQRT(4*pi/5)))
         es = MAX(es,ABS(ms(lm(lmax,2,0),k)+(1.0D0/3)*om*b0*r*SQRT(4*pi/5)))
         et = MAX(et,MAXVAL(ABS(mt(:,k))))
         DO m = 0,lmax
            DO l = m,lmax
               sh = MAX(sh,ABS(DCMPLX(0.0D0,m*om)*t(lm(lmax,l,m),k)))
               eh = MAX(eh,ABS(nh(lm(lmax,l,m),k)-DCMPLX(0.0D0,m*om)*t(lm(lmax,l,m),k)))
            END DO
         END DO
      END DO
      CALL report('u x B solid body in B0 z MQ (rel)',eq/sc,1.0D-9)
      CALL report('u x B solid body in B0 z MS (rel)',es/sc,1.0D-12)
      CALL report('u x B solid body in B0 z MT (abs)',et,1.0D-12)
      CALL report('u . grad T = Omega dT/dphi (rel)',eh/sh,1.0D-12)
      CALL gdev_free(C_LOC(w)); CALL gdev_free(C_LOC(z)); CALL gdev_free(C_LOC(pb))
      CALL gdev_free(C_LOC(tb)); CALL gdev_free(C_LOC(t)); CALL gdev_free(C_LOC(nq))
      CALL gdev_free(C_LOC(ns)); CALL gdev_free(C_LOC(nt)); CALL gdev_free(C_LOC(mq))
      CALL gdev_free(C_LOC(ms)); CALL gdev_free(C_LOC(mt)); CALL gdev_free(C_LOC(nh))
      CALL fcgram_destroy_plan(planfc)
      END SUBROUTINE caso_coupled

! Device copy of a spectral host array (registered by its address).
      SUBROUTINE dev_all(a)
      DOUBLE COMPLEX, TARGET :: a(:,:)
      CALL gdev_alloc(C_LOC(a),INT(16*SIZE(a),C_SIZE_T))
      END SUBROUTINE dev_all

! prueba_boundary (src/prueba_boundary.f90:58-128): lmax = 7, nz = 128,
! rint = 0.5, rext = 1.5, through the boundary module, at the levels
! LEEME.md:45 records (walls exact, dW/dr 1e-14, interior phi 3e-12,
! twice 2e-13, W vs W0 6e-8, fast == full, fast phi 3e-12).
      SUBROUTINE caso_boundary()
      USE boundary
      INTEGER, PARAMETER :: lmax = 7, nz = 128, C = 25, od = 5, n = nz - C
      INTEGER, PARAMETER :: nlm = (lmax+1)*(lmax+2)/2, nlat = 12, nphi = 24
      INTEGER, PARAMETER :: nf = 9
      TYPE(BCPLAN) :: bcp
      CHARACTER(len=15) :: bck(2)
      DOUBLE COMPLEX, TARGET :: a(nlm,nz,nf), z0(nlm,nz), w0(nlm,nz)
      DOUBLE PRECISION, TARGET :: zz(nz), d1(n,n)
      INTEGER(C_INT), TARGET :: ell(nlm)
      DOUBLE PRECISION :: dz, r1, r2, err(10), wm, dmax(2)
      DOUBLE COMPLEX :: cl
      INTEGER :: k, i, j
      INTEGER(C_SIZE_T) :: nb
      CHARACTER(len=256) :: tdir
! a(:,:,j): 1 f0, 2 tor, 3 fi, 4 pol, 5 pol2, 6 fs, 7 tor3, 8 fi3, 9 pol3
      CALL GET_ENVIRONMENT_VARIABLE('RD_TABLE_DIR',tdir)
      dz = 1.0D0/(n-1)
      CALL fcgram_create_plan(planfc,(/nlat,nphi,nz/),(/0,0,C/),(/0,0,od/), &
           (/1.0D0,1.0D0,dz/),TRIM(tdir),0,lmax,0.5D0)
      CALL rd_check(rd_plan_grid(planfc%h,C_NULL_PTR,C_NULL_PTR,C_NULL_PTR,C_LOC(zz), &
           C_NULL_PTR,C_LOC(ell),C_NULL_PTR),'grid')
      r1 = zz(1)
      r2 = zz(n)
      bck(1) = 'noslip'
      bck(2) = 'NoSlip '
      CALL setup_bc(bcp,planfc,bck,'v')
      DO k = 1,nz
         DO i = 1,nlm
            cl = DCMPLX(1.0D0+0.1D0*i,0.5D0-0.05D0*i)
            z0(i,k) = cl*(1.0D0+EXP(zz(k))*SIN(3*zz(k)))
            w0(i,k) = cl*(zz(k)-r1)**2*(zz(k)-r2)**2*EXP(zz(k))
         END DO
      END DO
      nb = INT(16*nlm*nz,C_SIZE_T)
      CALL gdev_alloc(C_LOC(w0),nb)
      DO j = 1,nf
         CALL gdev_alloc(C_LOC(a(1,1,j)),nb)
      END DO
      CALL gdev_update_to(C_LOC(w0),nb)
      CALL dlr(w0,a(1,1,1))
      CALL gdev_update_from(C_LOC(a(1,1,1)),nb)
! what a stage leaves on the wall rows and in l = 0
      a(:,:,2) = z0
      a(:,:,3) = a(:,:,1)
      a(:,1,3) = a(:,1,3)+1.0D0
      a(:,2,3) = a(:,2,3)-2.0D0
      a(:,n-1,3) = a(:,n-1,3)+3.0D0
      a(:,n,3) = a(:,n,3)-4.0D0
      DO i = 1,nlm
         IF (ell(i) .eq. 0) a(i,:,3) = (5.0D0,1.0D0)
      END DO
      a(:,:,6) = a(:,:,3)
      a(:,:,7) = z0
      a(:,:,8) = a(:,:,3)
      DO j = 2,8
         CALL gdev_update_to(C_LOC(a(1,1,j)),nb)
      END DO
      CALL v_imposebc(a(1,1,2),a(1,1,3),a(1,1,4))
      CALL v_imposebc(a(1,1,2),a(1,1,3),a(1,1,5))
      CALL v_imposebc_fast(a(1,1,7),a(1,1,8),a(1,1,9))
      DO j = 2,nf
         CALL gdev_update_from(C_LOC(a(1,1,j)),nb)
      END DO
      CALL derivr_mat(1,d1)
      err = 0
      err(1) = MAX(MAXVAL(ABS(a(:,1,2))),MAXVAL(ABS(a(:,n,2))))
      err(2) = MAX(MAXVAL(ABS(a(:,2:n-1,2)-z0(:,2:n-1))),MAXVAL(ABS(a(:,n+1:nz,2)-z0(:,n+1:nz))))
      DO i = 1,nlm
         IF (ell(i) .eq. 0) THEN
            err(5) = MAX(err(5),MAXVAL(ABS(a(i,:,3))),MAXVAL(ABS(a(i,:,4))))
            CYCLE
         ENDIF
         wm = MAXVAL(ABS(a(i,1:n,4)))
         err(3) = MAX(err(3),MAX(ABS(a(i,1,4)),ABS(a(i,n,4)))/wm)
         err(4) = MAX(err(4),MAX(ABS(SUM(d1(1,:)*a(i,1:n,4))),ABS(SUM(d1(n,:)*a(i,1:n,4))))*(r2-r1)/wm)
         err(6) = MAX(err(6),MAXVAL(ABS(a(i,3:n-2,3)-a(i,3:n-2,1)))/MAXVAL(ABS(a(i,3:n-2,1))))
         err(8) = MAX(err(8),MAXVAL(ABS(a(i,1:n,4)-w0(i,1:n)))/MAXVAL(ABS(w0(i,1:n))))
      END DO
      err(7) = MAXVAL(ABS(a(:,:,5)-a(:,:,4)))/MAXVAL(ABS(a(:,:,4)))
      err(9) = MAX(MAXVAL(ABS(a(:,:,9)-a(:,:,4))),MAXVAL(ABS(a(:,:,7)-a(:,:,2))))
      dmax(1) = MAXVAL(ABS(a(:,1:n,8)-a(:,1:n,3)))
      dmax(2) = MAXVAL(ABS(a(:,1:n,3)))
      err(10) = dmax(1)/dmax(2)
      CALL report('boundary: Z at the walls',err(1),0.0D0)
      CALL report('boundary: Z off the walls, change',err(2),0.0D0)
      CALL report('boundary: W at the walls (rel)',err(3),1.0D-13)
      CALL report('boundary: dW/dr at the walls (rel)',err(4),1.0D-12)
      CALL report('boundary: phi and W at l = 0',err(5),0.0D0)
      CALL report('boundary: interior phi change',err(6),3.0D-11)
      CALL report('boundary: applied twice, W change',err(7),2.0D-12)
      CALL report('boundary: W vs W0 (truncation 6e-8)',err(8),6.0D-7)
      CALL report('boundary: fast, W and Z vs full',err(9),0.0D0)
      CALL report('boundary: fast, phi vs full',err(10),3.0D-11)
      CALL gdev_free(C_LOC(w0))
      DO j = 1,nf
         CALL gdev_free(C_LOC(a(1,1,j)))
      END DO
      CALL fcgram_destroy_plan(planfc)
      END SUBROUTINE caso_boundary

! grid_init and box_init from a namelist file, as a reference program
! starts (pseudospec_mod.f90:41-69, 126-236): the 3/2 rule, the Gauss
! nodes of the plan, the radial grid ending at rext on z(nz-Cz), kz
! in FFT order, (l,m) in the m-major order.
      SUBROUTINE caso_box()
      USE grid
      USE kes
      USE boxsize
      TYPE(FCPLAN) :: plan
      DOUBLE PRECISION, TARGET :: ct(24)
      DOUBLE PRECISION :: e1, e2, e3
      INTEGER :: i, l, m, bad
      OPEN(7,file='prueba_rd.nml',action='write')
      WRITE(7,'(A)') '&grid'
      WRITE(7,'(A)') ' lmax = 15, nz = 89, Cz = 25, oz = 5'
      WRITE(7,'(A)') '/'
      WRITE(7,'(A)') '&boxparams'
      WRITE(7,'(A)') ' rint = 0.35, rext = 1.0'
      WRITE(7,'(A)') '/'
      CLOSE(7)
      CALL grid_init('prueba_rd.nml')
      CALL box_init('prueba_rd.nml')
      bad = 0
      IF (nlm .ne. 136 .OR. nlat .ne. 24 .OR. nphi .ne. 48 .OR. Cz .ne. 25 .OR. oz .ne. 5) bad = 1
      CALL fcgram_create_plan(plan,(/nlat,nphi,nz/),(/0,0,0/),(/0,0,0/), &
           (/1.0D0,1.0D0,1.0D0/),'.',0,lmax)
      CALL rd_check(rd_plan_grid(plan%h,C_LOC(ct),C_NULL_PTR,C_NULL_PTR,C_NULL_PTR, &
           C_NULL_PTR,C_NULL_PTR,C_NULL_PTR),'grid')
      e1 = MAXVAL(ABS(COS(theta)-ct))
      e2 = ABS(z(nz-Cz)-rext) + ABS(z(1)-rint) + ABS(kz(2)-Dkz) + ABS(kz(nz)+Dkz) + &
           ABS(kz(45)-44*Dkz) + ABS(kz(46)+44*Dkz) + ABS(phi(2)-2*pi/nphi)
      DO m = 0,lmax
         DO l = m,lmax
            i = lm(lmax,l,m)
            IF (ell(i) .ne. l .OR. emm(i) .ne. m) bad = bad + 1
         END DO
      END DO
      e3 = ABS(Dkz - 2*pi/((rext-rint)/(nz-Cz-1)*nz))
      CALL report('box_init theta vs plan nodes',e1,1.0D-15)
      CALL report('box_init z, kz, phi',e2,1.0D-13)
      CALL report('box_init Dkz',e3,1.0D-12)
      CALL report('grid_init sizes, box_init ell/emm (bad)',DBLE(bad),0.0D0)
      CALL fcgram_destroy_plan(plan)
      END SUBROUTINE caso_box

! load_neumann_tables, neumann_reconstruct and robin_reconstruct
! (fcgram_mod.f90:267-650) on the radial rows: the wall value of a
! degree-4 profile from its outward normal derivative.
      SUBROUTINE caso_neumann()
      INTEGER, PARAMETER :: lmax = 7, nr = 40, C = 25, od = 5, nz = nr + C
      INTEGER, PARAMETER :: nlat = 12, nphi = 24, nlm = (lmax+1)*(lmax+2)/2
      DOUBLE PRECISION, PARAMETER :: ri = 0.5D0, ro = 1.5D0
      DOUBLE COMPLEX, TARGET :: f(nlm,nz), g(nlm,nz)
      DOUBLE PRECISION, TARGET :: a(nlm,1)
      DOUBLE PRECISION :: dz, r, e5, e6, e2, er
      DOUBLE COMPLEX :: c0, c1, c2, c3, c4
      INTEGER :: j, k
      CHARACTER(len=256) :: tdir
      CALL GET_ENVIRONMENT_VARIABLE('RD_TABLE_DIR',tdir)
      dz = (ro-ri)/(nr-1)
      CALL fcgram_create_plan(planfc,(/nlat,nphi,nz/),(/0,0,C/),(/0,0,od/), &
           (/1.0D0,1.0D0,dz/),TRIM(tdir),0,lmax,ri)
      CALL load_neumann_tables(planfc%z,TRIM(tdir),1)
      CALL load_neumann_tables(planfc%z,TRIM(tdir),2)
      CALL load_neumann_tables(planfc%z,TRIM(tdir),1)     ! warns, keeps the table
      c0 = (0.3D0,0.1D0); c1 = (-1.2D0,0.4D0); c2 = (0.7D0,-0.2D0); c3 = (0.25D0,0.5D0); c4 = (-0.4D0,0.1D0)
      DO k = 1,nz
         r = ri+(k-1)*dz
         DO j = 1,nlm
            f(j,k) = j*(c0+c1*r+c2*r**2+c3*r**3+c4*r**4)
         END DO
      END DO
      DO j = 1,nlm
         a(j,1) = 0.5D0+0.01D0*j
      END DO
      CALL gdev_alloc(C_LOC(g),INT(16*SIZE(g),C_SIZE_T))
      e5 = 0; e6 = 0; e2 = 0; er = 0
      ! boundary 5, first derivative: g = -df/dr at r = ri
      g = f
      r = ri
      g(:,1) = -(/(j*(c1+2*c2*r+3*c3*r**2+4*c4*r**3), j=1,nlm)/)
      CALL gdev_update_to(C_LOC(g),INT(16*SIZE(g),C_SIZE_T))
      CALL neumann_reconstruct(planfc,g,5,1)
      CALL gdev_update_from(C_LOC(g),INT(16*SIZE(g),C_SIZE_T))
      e5 = MAXVAL(ABS(g(:,1)-f(:,1)))/MAXVAL(ABS(f(:,1)))
      ! boundary 6, first derivative: g = df/dr at r = ro
      g = f
      r = ro
      g(:,nr) = (/(j*(c1+2*c2*r+3*c3*r**2+4*c4*r**3), j=1,nlm)/)
      CALL gdev_update_to(C_LOC(g),INT(16*SIZE(g),C_SIZE_T))
      CALL neumann_reconstruct(planfc,g,6,1)
      CALL gdev_update_from(C_LOC(g),INT(16*SIZE(g),C_SIZE_T))
      e6 = MAXVAL(ABS(g(:,nr)-f(:,nr)))/MAXVAL(ABS(f(:,nr)))
      ! boundary 6, second derivative
      g = f
      g(:,nr) = (/(j*(2*c2+6*c3*r+12*c4*r**2), j=1,nlm)/)
      CALL gdev_update_to(C_LOC(g),INT(16*SIZE(g),C_SIZE_T))
      CALL neumann_reconstruct(planfc,g,6,2)
      CALL gdev_update_from(C_LOC(g),INT(16*SIZE(g),C_SIZE_T))
      e2 = MAXVAL(ABS(g(:,nr)-f(:,nr)))/MAXVAL(ABS(f(:,nr)))
      ! Robin at boundary 6: g = f' + a f
      g = f
      g(:,nr) = (/(j*(c1+2*c2*r+3*c3*r**2+4*c4*r**3), j=1,nlm)/) + a(:,1)*f(:,nr)
      CALL gdev_update_to(C_LOC(g),INT(16*SIZE(g),C_SIZE_T))
      CALL robin_reconstruct(planfc,g,6,a)
      CALL gdev_update_from(C_LOC(g),INT(16*SIZE(g),C_SIZE_T))
      er = MAXVAL(ABS(g(:,nr)-f(:,nr)))/MAXVAL(ABS(f(:,nr)))
      CALL report('neumann_reconstruct z=rint order 1',e5,1.0D-12)
      CALL report('neumann_reconstruct z=rext order 1',e6,1.0D-12)
      CALL report('neumann_reconstruct z=rext order 2',e2,1.0D-12)
      CALL report('robin_reconstruct z=rext',er,1.0D-12)
      CALL gdev_free(C_LOC(g))
      CALL fcgram_destroy_plan(planfc)
      END SUBROUTINE caso_neumann

! fftp3d_redist_k_to_lm / _lm_to_k (sphp_sca.F90:629-760): the one-rank
! exchange copies the block; gdev_lookup finds registered arrays only.
      SUBROUTINE caso_redist()
      INTEGER, PARAMETER :: lmax = 7, nz = 6, nlat = 12, nphi = 24, nlm = (lmax+1)*(lmax+2)/2
      TYPE(FCPLAN) :: plan
      DOUBLE COMPLEX, TARGET :: a(nlm,nz), b(nlm,nz), c(nlm,nz)
      DOUBLE PRECISION, TARGET :: e, unreg(4)
      INTEGER :: j, k, ia
      CALL fcgram_create_plan(plan,(/nlat,nphi,nz/),(/0,0,0/),(/0,0,0/), &
           (/1.0D0,1.0D0,1.0D0/),'.',0,lmax)
      DO k = 1,nz
         DO j = 1,nlm
            a(j,k) = DCMPLX(j+0.5D0*k,-k*0.25D0*j)
         END DO
      END DO
      b = 0; c = 0
      CALL gdev_alloc(C_LOC(a),INT(16*SIZE(a),C_SIZE_T))
      CALL gdev_alloc(C_LOC(b),INT(16*SIZE(b),C_SIZE_T))
      CALL gdev_alloc(C_LOC(c),INT(16*SIZE(c),C_SIZE_T))
      CALL gdev_update_to(C_LOC(a),INT(16*SIZE(a),C_SIZE_T))
      CALL fftp3d_redist_k_to_lm(plan,a,b,0)
      CALL fftp3d_redist_lm_to_k(plan,b,c,0)
      CALL gdev_update_from(C_LOC(b),INT(16*SIZE(b),C_SIZE_T))
      CALL gdev_update_from(C_LOC(c),INT(16*SIZE(c),C_SIZE_T))
      e = MAXVAL(ABS(b-a)) + MAXVAL(ABS(c-a))
      CALL report('fftp3d_redist k->lm->k (copy)',e,0.0D0)
      ia = gdev_lookup(C_LOC(a))
      e = 0
      IF (ia .lt. 1 .OR. gdev_lookup(C_LOC(unreg)) .ne. 0) e = 1
      CALL report('gdev_lookup registered / unregistered',e,0.0D0)
      CALL gdev_free(C_LOC(a)); CALL gdev_free(C_LOC(b)); CALL gdev_free(C_LOC(c))
      IF (gdev_lookup(C_LOC(a)) .ne. 0) e = 1
      CALL report('gdev_lookup after gdev_free',e,0.0D0)
      CALL fcgram_destroy_plan(plan)
      END SUBROUTINE caso_redist

      END PROGRAM prueba_rd
