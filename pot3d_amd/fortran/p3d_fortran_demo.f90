!#######################################################################
! A slim potfld (src/pot3d.F90:3866-3976) over the C ABI: uniform meshes as
! set_global_mesh builds them (src/pot3d.F90:2583-2706), a Br map at r0,
! the PCG solve with ILU0 on the GPU, phi with boundary values, and the
! energy diagnostic.  Prints one line the Python test checks.
!#######################################################################
program p3d_fortran_demo
      use pot3d_hip
      use, intrinsic :: iso_c_binding
      implicit none
      integer, parameter :: nr=17, nt=19, np=28
      real(c_double), parameter :: pi=3.1415926535897932_c_double
      real(c_double) :: r(nr-1),dr(nr-1),rh(nr),drh(nr)
      real(c_double) :: t(nt-1),st(nt-1),dt(nt-1),th(nt),sth(nt),dth(nt)
      real(c_double) :: p(np),dp(np),ph(np),dph(np)
      real(c_double), allocatable :: br0(:,:),x(:),phi(:)
      real(c_double) :: w(3),epsn
      type(c_ptr) :: h
      integer(c_int) :: ierr,ncg
      integer :: i,j,k,n
!
      do i=1,nr-1
        r(i)=1._c_double+1.5_c_double*(i-1)/(nr-2)
      enddo
      do i=2,nr-1
        rh(i)=.5_c_double*(r(i)+r(i-1))
        drh(i)=r(i)-r(i-1)
      enddo
      rh(1)=rh(2)-drh(2); rh(nr)=rh(nr-1)+drh(nr-1)
      drh(1)=drh(2); drh(nr)=drh(nr-1)
      do i=1,nr-1
        dr(i)=rh(i+1)-rh(i)
      enddo
      do j=1,nt-1
        t(j)=pi*(j-1)/(nt-2)
      enddo
      do j=2,nt-1
        th(j)=.5_c_double*(t(j)+t(j-1))
        dth(j)=t(j)-t(j-1)
      enddo
      th(1)=th(2)-dth(2); th(nt)=th(nt-1)+dth(nt-1)
      dth(1)=dth(2); dth(nt)=dth(nt-1)
      do j=1,nt-1
        dt(j)=th(j+1)-th(j)
      enddo
      st=sin(t); sth=sin(th)
      do k=1,np-1
        p(k)=2._c_double*pi*(k-1)/(np-2)
      enddo
      p(np)=p(2)+2._c_double*pi
      do k=2,np
        ph(k)=.5_c_double*(p(k)+p(k-1))
        dph(k)=p(k)-p(k-1)
      enddo
      ph(1)=ph(np-1)-2._c_double*pi; dph(1)=dph(np-1)
      do k=1,np-1
        dp(k)=ph(k+1)-ph(k)
      enddo
      dp(np)=dp(2); dph(np)=dph(2); dp(1)=dp(np-1)
!
      allocate (br0(nt,np))
      do k=1,np
        do j=1,nt
          br0(j,k)=cos(th(j))
        enddo
      enddo
      n=(nr-2)*(nt-2)*(np-2)
      allocate (x(n),phi(nr*nt*np))
      x=0.
!
      ierr=p3d_create(h,nr,nt,np,r,dr,rh,drh,st,dt,sth,dth,dp,dph, &
                      -1._c_double,2_c_int)
      if (ierr.ne.P3D_OK) stop 'p3d_create failed'
      ierr=p3d_set_br0(h,br0)
      ierr=p3d_solve(h,x,c_null_ptr,1.e-9_c_double,10000_c_int,0_c_int, &
                     ncg,epsn)
      if (ierr.ne.P3D_OK) stop 'p3d_solve failed'
      ierr=p3d_fill_phi(h,c_null_ptr,phi)
      ierr=p3d_energy(h,c_null_ptr,w,c_null_ptr,c_null_ptr,c_null_ptr)
      write (*,'(a,i6,3es24.16)') 'FORTRAN_DEMO ',ncg,w
      ierr=p3d_destroy(h)
end program
