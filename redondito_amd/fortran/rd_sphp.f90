!=================================================================
! External routines of the reference's sphp / pseudo / boundary path
! with their argument lists, over the C-ABI of librdsphp.  Arrays are
! the caller's host arrays; the work is done on their registered device
! copies (gdev_alloc / gdev_update_to), as in the reference's GPU build
! with gdev_active set.  One rank holds the whole domain: physical arrays
! are (nlat,nphi,nz), spectral ones (nlm,nz).  comm is accepted and not
! used.  Each routine waits for its device work before returning.
!=================================================================

! src/sphp/sphp_sca.F90:273-329
      SUBROUTINE fftp2d_real_to_complex_xy(plan,in,out,comm)
      USE fcgram
      IMPLICIT NONE
      TYPE(FCPLAN), INTENT(IN) :: plan
      DOUBLE PRECISION, INTENT(INOUT), TARGET :: in(*)
      DOUBLE COMPLEX, INTENT(OUT), TARGET :: out(*)
      TYPE(*), INTENT(IN) :: comm
      CALL rd_check(rd_sht_analysis(plan%h,INT(plan%z%n,C_INT),rd_dptr(C_LOC(in),'in'), &
           rd_dptr(C_LOC(out),'out')),'fftp2d_real_to_complex_xy')
      CALL fc_sync(plan)
      END SUBROUTINE fftp2d_real_to_complex_xy

! src/sphp/sphp_sca.F90:436-492
      SUBROUTINE fftp2d_complex_to_real_xy(plan,in,out,comm)
      USE fcgram
      IMPLICIT NONE
      TYPE(FCPLAN), INTENT(IN) :: plan
      DOUBLE COMPLEX, INTENT(IN), TARGET :: in(*)
      DOUBLE PRECISION, INTENT(OUT), TARGET :: out(*)
      TYPE(*), INTENT(IN) :: comm
      CALL rd_check(rd_sht_synthesis(plan%h,INT(plan%z%n,C_INT),rd_dptr(C_LOC(in),'in'), &
           rd_dptr(C_LOC(out),'out')),'fftp2d_complex_to_real_xy')
      CALL fc_sync(plan)
      END SUBROUTINE fftp2d_complex_to_real_xy

! src/sphp/sphp_vec.F90:49-106
      SUBROUTINE fftp2d_vec_to_qst(plan,vr,vt,vp,q,s,t,comm)
      USE fcgram
      IMPLICIT NONE
      TYPE(FCPLAN), INTENT(IN) :: plan
      DOUBLE PRECISION, INTENT(INOUT), TARGET :: vr(*),vt(*),vp(*)
      DOUBLE COMPLEX, INTENT(OUT), TARGET :: q(*),s(*),t(*)
      TYPE(*), INTENT(IN) :: comm
      CALL rd_check(rd_vsht_analysis(plan%h,INT(plan%z%n,C_INT),rd_dptr(C_LOC(vr),'vr'), &
           rd_dptr(C_LOC(vt),'vt'),rd_dptr(C_LOC(vp),'vp'),rd_dptr(C_LOC(q),'q'), &
           rd_dptr(C_LOC(s),'s'),rd_dptr(C_LOC(t),'t')),'fftp2d_vec_to_qst')
      CALL fc_sync(plan)
      END SUBROUTINE fftp2d_vec_to_qst

! src/sphp/sphp_vec.F90:109-165
      SUBROUTINE fftp2d_qst_to_vec(plan,q,s,t,vr,vt,vp,comm)
      USE fcgram
      IMPLICIT NONE
      TYPE(FCPLAN), INTENT(IN) :: plan
      DOUBLE COMPLEX, INTENT(IN), TARGET :: q(*),s(*),t(*)
      DOUBLE PRECISION, INTENT(OUT), TARGET :: vr(*),vt(*),vp(*)
      TYPE(*), INTENT(IN) :: comm
      CALL rd_check(rd_vsht_synthesis(plan%h,INT(plan%z%n,C_INT),rd_dptr(C_LOC(q),'q'), &
           rd_dptr(C_LOC(s),'s'),rd_dptr(C_LOC(t),'t'),rd_dptr(C_LOC(vr),'vr'), &
           rd_dptr(C_LOC(vt),'vt'),rd_dptr(C_LOC(vp),'vp')),'fftp2d_qst_to_vec')
      CALL fc_sync(plan)
      END SUBROUTINE fftp2d_qst_to_vec

! src/sphp/sphp_sca.F90:332-397
      SUBROUTINE fftp1d_real_to_complex_z(plan,inout,comm)
      USE fcgram
      IMPLICIT NONE
      TYPE(FCPLAN), INTENT(IN) :: plan
      DOUBLE COMPLEX, INTENT(INOUT), TARGET :: inout(*)
      TYPE(*), INTENT(IN) :: comm
      CALL rd_check(rd_z_forward(plan%h,rd_dptr(C_LOC(inout),'inout')),'fftp1d_real_to_complex_z')
      CALL fc_sync(plan)
      END SUBROUTINE fftp1d_real_to_complex_z

! src/sphp/sphp_sca.F90:495-530
      SUBROUTINE fftp1d_complex_to_real_z(plan,inout,comm)
      USE fcgram
      IMPLICIT NONE
      TYPE(FCPLAN), INTENT(IN) :: plan
      DOUBLE COMPLEX, INTENT(INOUT), TARGET :: inout(*)
      TYPE(*), INTENT(IN) :: comm
      CALL rd_check(rd_z_inverse(plan%h,rd_dptr(C_LOC(inout),'inout')),'fftp1d_complex_to_real_z')
      CALL fc_sync(plan)
      END SUBROUTINE fftp1d_complex_to_real_z

! src/sphp/sphp_sca.F90:237-270
      SUBROUTINE fftp3d_real_to_complex(plan,in,out,comm)
      USE fcgram
      IMPLICIT NONE
      TYPE(FCPLAN), INTENT(IN) :: plan
      DOUBLE PRECISION, INTENT(INOUT), TARGET :: in(*)
      DOUBLE COMPLEX, INTENT(OUT), TARGET :: out(*)
      TYPE(*), INTENT(IN) :: comm
      CALL rd_check(rd_3d_forward(plan%h,rd_dptr(C_LOC(in),'in'),rd_dptr(C_LOC(out),'out')), &
           'fftp3d_real_to_complex')
      CALL fc_sync(plan)
      END SUBROUTINE fftp3d_real_to_complex

! src/sphp/sphp_sca.F90:400-433
      SUBROUTINE fftp3d_complex_to_real(plan,in,out,comm)
      USE fcgram
      IMPLICIT NONE
      TYPE(FCPLAN), INTENT(IN) :: plan
      DOUBLE COMPLEX, INTENT(INOUT), TARGET :: in(*)
      DOUBLE PRECISION, INTENT(OUT), TARGET :: out(*)
      TYPE(*), INTENT(IN) :: comm
      CALL rd_check(rd_3d_inverse(plan%h,rd_dptr(C_LOC(in),'in'),rd_dptr(C_LOC(out),'out')), &
           'fftp3d_complex_to_real')
      CALL fc_sync(plan)
      END SUBROUTINE fftp3d_complex_to_real

! src/sphp/sphp_sca.F90:629-693: planes -> lm pencils; one rank holds
! every plane and lm, so the exchange is a copy of the (nlm,nz) block
      SUBROUTINE fftp3d_redist_k_to_lm(plan,in,out,comm)
      USE fcgram
      IMPLICIT NONE
      TYPE(FCPLAN), INTENT(IN) :: plan
      DOUBLE COMPLEX, INTENT(IN), TARGET :: in(*)
      DOUBLE COMPLEX, INTENT(OUT), TARGET :: out(*)
      TYPE(*), INTENT(IN) :: comm
      CALL rd_check(rd_redist_k_to_lm(plan%h,INT(plan%z%n,C_INT),rd_dptr(C_LOC(in),'in'), &
           rd_dptr(C_LOC(out),'out')),'fftp3d_redist_k_to_lm')
      CALL fc_sync(plan)
      END SUBROUTINE fftp3d_redist_k_to_lm

! src/sphp/sphp_sca.F90:696-760: lm pencils -> planes (a copy on one rank)
      SUBROUTINE fftp3d_redist_lm_to_k(plan,in,out,comm)
      USE fcgram
      IMPLICIT NONE
      TYPE(FCPLAN), INTENT(IN) :: plan
      DOUBLE COMPLEX, INTENT(IN), TARGET :: in(*)
      DOUBLE COMPLEX, INTENT(OUT), TARGET :: out(*)
      TYPE(*), INTENT(IN) :: comm
      CALL rd_check(rd_redist_lm_to_k(plan%h,INT(plan%z%n,C_INT),rd_dptr(C_LOC(in),'in'), &
           rd_dptr(C_LOC(out),'out')),'fftp3d_redist_lm_to_k')
      CALL fc_sync(plan)
      END SUBROUTINE fftp3d_redist_lm_to_k

!-----------------------------------------------------------------
! src/pseudo/pseudospec_hd.f90, on the global plan planfc (module fft)
!-----------------------------------------------------------------
      SUBROUTINE derivr(a,b,p,esp)
      USE fft
      IMPLICIT NONE
      DOUBLE COMPLEX, INTENT(IN), TARGET :: a(*)
      DOUBLE COMPLEX, INTENT(OUT), TARGET :: b(*)
      INTEGER, INTENT(IN) :: p
      CHARACTER(len=1), INTENT(IN) :: esp
      IF (esp .ne. 'r' .AND. esp .ne. 'k') THEN
         PRINT*, "[ERROR] Unknown space ", esp, " in derivr. Aborting..."
         STOP
      ENDIF
      CALL rd_check(rd_derivr(planfc%h,INT(p,C_INT),esp,rd_dptr(C_LOC(a),'a'), &
           rd_dptr(C_LOC(b),'b')),'derivr')
      CALL fc_sync(planfc)
      END SUBROUTINE derivr

      SUBROUTINE laplang(a,b)
      USE fft
      IMPLICIT NONE
      DOUBLE COMPLEX, INTENT(IN), TARGET :: a(*)
      DOUBLE COMPLEX, INTENT(OUT), TARGET :: b(*)
      CALL rd_check(rd_laplang(planfc%h,rd_dptr(C_LOC(a),'a'),rd_dptr(C_LOC(b),'b')),'laplang')
      CALL fc_sync(planfc)
      END SUBROUTINE laplang

      SUBROUTINE multr(a,b,p)
      USE fft
      IMPLICIT NONE
      DOUBLE COMPLEX, INTENT(IN), TARGET :: a(*)
      DOUBLE COMPLEX, INTENT(OUT), TARGET :: b(*)
      INTEGER, INTENT(IN) :: p
      CALL rd_check(rd_multr(planfc%h,INT(p,C_INT),rd_dptr(C_LOC(a),'a'),rd_dptr(C_LOC(b),'b')),'multr')
      CALL fc_sync(planfc)
      END SUBROUTINE multr

      SUBROUTINE dlr(a,b)
      USE fft
      IMPLICIT NONE
      DOUBLE COMPLEX, INTENT(IN), TARGET :: a(*)
      DOUBLE COMPLEX, INTENT(OUT), TARGET :: b(*)
      CALL rd_check(rd_dlr(planfc%h,rd_dptr(C_LOC(a),'a'),rd_dptr(C_LOC(b),'b')),'dlr')
      CALL fc_sync(planfc)
      END SUBROUTINE dlr

      SUBROUTINE saxpby_c(z,x,a,y,b)
      USE fft
      IMPLICIT NONE
      DOUBLE COMPLEX, INTENT(INOUT), TARGET :: z(*)
      DOUBLE COMPLEX, INTENT(IN), TARGET :: x(*),y(*)
      DOUBLE PRECISION, INTENT(IN) :: a,b
      CALL rd_check(rd_saxpby_c(planfc%h,rd_dptr(C_LOC(z),'z'),rd_dptr(C_LOC(x),'x'), &
           REAL(a,C_DOUBLE),rd_dptr(C_LOC(y),'y'),REAL(b,C_DOUBLE)),'saxpby_c')
      CALL fc_sync(planfc)
      END SUBROUTINE saxpby_c

      SUBROUTINE saxpby_r(z,x,a,y,b)
      USE fft
      IMPLICIT NONE
      DOUBLE PRECISION, INTENT(INOUT), TARGET :: z(*)
      DOUBLE PRECISION, INTENT(IN), TARGET :: x(*),y(*)
      DOUBLE PRECISION, INTENT(IN) :: a,b
      CALL rd_check(rd_saxpby_r(planfc%h,INT(planfc%z%n,C_INT),rd_dptr(C_LOC(z),'z'), &
           rd_dptr(C_LOC(x),'x'),REAL(a,C_DOUBLE),rd_dptr(C_LOC(y),'y'),REAL(b,C_DOUBLE)),'saxpby_r')
      CALL fc_sync(planfc)
      END SUBROUTINE saxpby_r

      SUBROUTINE prodvec(ar,at,ap,br,bt,bp,cr,ct,cp)
      USE fft
      IMPLICIT NONE
      DOUBLE PRECISION, INTENT(IN), TARGET :: ar(*),at(*),ap(*),br(*),bt(*),bp(*)
      DOUBLE PRECISION, INTENT(OUT), TARGET :: cr(*),ct(*),cp(*)
      CALL rd_check(rd_prodvec(planfc%h,INT(planfc%z%n,C_INT),rd_dptr(C_LOC(ar),'ar'), &
           rd_dptr(C_LOC(at),'at'),rd_dptr(C_LOC(ap),'ap'),rd_dptr(C_LOC(br),'br'), &
           rd_dptr(C_LOC(bt),'bt'),rd_dptr(C_LOC(bp),'bp'),rd_dptr(C_LOC(cr),'cr'), &
           rd_dptr(C_LOC(ct),'ct'),rd_dptr(C_LOC(cp),'cp')),'prodvec')
      CALL fc_sync(planfc)
      END SUBROUTINE prodvec

      SUBROUTINE fc_filter(a,esp)
      USE fft
      IMPLICIT NONE
      DOUBLE COMPLEX, INTENT(INOUT), TARGET :: a(*)
      CHARACTER(len=1), INTENT(IN) :: esp
      IF (esp .ne. 'r' .AND. esp .ne. 'k') THEN
         PRINT*, "[ERROR] Unknown space ", esp, " in fc_filter. Aborting..."
         STOP
      ENDIF
      CALL rd_check(rd_fc_filter(planfc%h,esp,rd_dptr(C_LOC(a),'a')),'fc_filter')
      CALL fc_sync(planfc)
      END SUBROUTINE fc_filter

      SUBROUTINE copy(a,b)
      USE fft
      IMPLICIT NONE
      DOUBLE COMPLEX, INTENT(IN), TARGET :: a(*)
      DOUBLE COMPLEX, INTENT(OUT), TARGET :: b(*)
      CALL rd_check(rd_copy(planfc%h,rd_dptr(C_LOC(a),'a'),rd_dptr(C_LOC(b),'b')),'copy')
      CALL fc_sync(planfc)
      END SUBROUTINE copy

! The composed nonlinear term N = u x omega (no reference routine):
! W, Z in (lm,r) -> NQ, NS, NT in (lm,r); prec 0 = FP64.
      SUBROUTINE nonlinear_uxw(w,z,nq,ns,nt,prec)
      USE fft
      IMPLICIT NONE
      DOUBLE COMPLEX, INTENT(IN), TARGET :: w(*),z(*)
      DOUBLE COMPLEX, INTENT(OUT), TARGET :: nq(*),ns(*),nt(*)
      INTEGER, INTENT(IN) :: prec
      CALL rd_check(rd_nonlinear_uxw(planfc%h,INT(prec,C_INT),rd_dptr(C_LOC(w),'w'), &
           rd_dptr(C_LOC(z),'z'),rd_dptr(C_LOC(nq),'nq'),rd_dptr(C_LOC(ns),'ns'), &
           rd_dptr(C_LOC(nt),'nt')),'nonlinear_uxw')
      CALL fc_sync(planfc)
      END SUBROUTINE nonlinear_uxw

! One nonlinear step of the coupled state (no reference routine): velocity
! W, Z, magnetic field Pb, Tb (poloidal, toroidal) and temperature T in
! (lm,r) -> u x omega (NQ,NS,NT), u x B (MQ,MS,MT), u . grad T (NH).
      SUBROUTINE nonlinear_coupled(w,z,pb,tb,t,nq,ns,nt,mq,ms,mt,nh,prec)
      USE fft
      IMPLICIT NONE
      DOUBLE COMPLEX, INTENT(IN), TARGET :: w(*),z(*),pb(*),tb(*),t(*)
      DOUBLE COMPLEX, INTENT(OUT), TARGET :: nq(*),ns(*),nt(*),mq(*),ms(*),mt(*),nh(*)
      INTEGER, INTENT(IN) :: prec
      CALL rd_check(rd_nonlinear_coupled(planfc%h,INT(prec,C_INT),rd_dptr(C_LOC(w),'w'), &
           rd_dptr(C_LOC(z),'z'),rd_dptr(C_LOC(pb),'pb'),rd_dptr(C_LOC(tb),'tb'), &
           rd_dptr(C_LOC(t),'t'),rd_dptr(C_LOC(nq),'nq'),rd_dptr(C_LOC(ns),'ns'), &
           rd_dptr(C_LOC(nt),'nt'),rd_dptr(C_LOC(mq),'mq'),rd_dptr(C_LOC(ms),'ms'), &
           rd_dptr(C_LOC(mt),'mt'),rd_dptr(C_LOC(nh),'nh')),'nonlinear_coupled')
      CALL fc_sync(planfc)
      END SUBROUTINE nonlinear_coupled
