!#######################################################################
! pot3d_hip: iso_c_binding interfaces to libpot3d_hip.so (the gfx950 PCG
! solve), for POT3D's Fortran driver.  Declarations mirror
! include/pot3d_hip.h.  The three legacy symbols keep the exact
! interfaces of POT3D's external_lu_solver_interface
! (src/pot3d.F90:581-609), so an unmodified driver built with -DHIPSPARSE
! links against this library instead of src/lusol_hipsparse.c.
!#######################################################################
module pot3d_hip
      use, intrinsic :: iso_c_binding
      implicit none
!
      integer(c_int), parameter :: P3D_OK=0, P3D_EINVAL=1, P3D_ENOMEM=2, &
                                   P3D_EHIP=3, P3D_ENOCONV=4,           &
                                   P3D_ENONFINITE=5, P3D_ESTATE=6
      integer(c_int), parameter :: P3D_OPT_STOP=1, P3D_OPT_MIXED=2,     &
                                   P3D_OPT_INNER_TOL=3, P3D_OPT_CHUNK=4,&
                                   P3D_OPT_GRAPH=5, P3D_OPT_SWEEP=6,    &
                                   P3D_OPT_VERBOSE=7,                   &
                                   P3D_OPT_GUESS_CHECK=8,               &
                                   P3D_OPT_TILE_K=9, P3D_OPT_TILE_WAVES=10, &
                                   P3D_OPT_ILU_BLOCKS_R=11,             &
                                   P3D_OPT_ILU_BLOCKS_T=12,             &
                                   P3D_OPT_ILU_BLOCKS_P=13,             &
                                   P3D_OPT_TILE_STAGE=14
!
      interface
! ****** Whole-solve ABI.
        integer(c_int) function p3d_create(h,nr,nt,np,r,dr,rh,drh,st,dt, &
                                           sth,dth,dp,dph,pm_r1,ifprec)  &
                       bind(c,name="p3d_create")
          import :: c_int, c_double, c_ptr
          type(c_ptr) :: h
          integer(c_int), value :: nr,nt,np,ifprec
          real(c_double) :: r(*),dr(*),rh(*),drh(*),st(*),dt(*)
          real(c_double) :: sth(*),dth(*),dp(*),dph(*)
          real(c_double), value :: pm_r1
        end function
        integer(c_int) function p3d_destroy(h) bind(c,name="p3d_destroy")
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function
        integer(c_int) function p3d_set_option(h,key,val) &
                       bind(c,name="p3d_set_option")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          integer(c_int), value :: key
          real(c_double), value :: val
        end function
        integer(c_int) function p3d_set_br0(h,br0) bind(c,name="p3d_set_br0")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double) :: br0(*)
        end function
        integer(c_int) function p3d_solve(h,x,rhs,epscg,ncgmax,ncghist, &
                                          ncg,epsn) bind(c,name="p3d_solve")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h, rhs
          real(c_double) :: x(*)
          real(c_double), value :: epscg
          integer(c_int), value :: ncgmax,ncghist
          integer(c_int) :: ncg
          real(c_double) :: epsn
        end function
        integer(c_int) function p3d_fill_phi(h,x,phi) &
                       bind(c,name="p3d_fill_phi")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h, x
          real(c_double) :: phi(*)
        end function
        integer(c_int) function p3d_energy(h,phi,w,br,bt,bp) &
                       bind(c,name="p3d_energy")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h, phi, br, bt, bp
          real(c_double) :: w(3)
        end function
! ****** Learned warm start (no counterpart in the reference).
        integer(c_int) function p3d_nn_create(net,nr,nt,np,rh,seed) &
                       bind(c,name="p3d_nn_create")
          import :: c_int, c_ptr, c_double
          type(c_ptr) :: net
          integer(c_int), value :: nr,nt,np,seed
          real(c_double) :: rh(*)
        end function
        integer(c_int) function p3d_nn_destroy(net) bind(c,name="p3d_nn_destroy")
          import :: c_int, c_ptr
          type(c_ptr), value :: net
        end function
        integer(c_int) function p3d_nn_set_scales(net,bscale,phiscale) &
                       bind(c,name="p3d_nn_set_scales")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: net
          real(c_double), value :: bscale,phiscale
        end function
        integer(c_int) function p3d_nn_add_sample(net,br0,x) &
                       bind(c,name="p3d_nn_add_sample")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: net
          real(c_double) :: br0(*),x(*)
        end function
        integer(c_int) function p3d_nn_train(net,nsteps,lr,seed,losses) &
                       bind(c,name="p3d_nn_train")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: net, losses
          integer(c_int), value :: nsteps,seed
          real(c_double), value :: lr
        end function
        integer(c_int) function p3d_nn_predict(net,br0,x,x_dev) &
                       bind(c,name="p3d_nn_predict")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: net, x_dev
          real(c_double) :: br0(*),x(*)
        end function
! ****** Legacy external LU solver ABI (src/pot3d.F90:585-603).
        subroutine load_external_lu_solver(csr_a,csr_ai,csr_aj,n,m) &
                   bind(c,name="load_external_lu_solver")
          import :: c_int, c_ptr
          integer(c_int), value :: n,m
          type(c_ptr), value :: csr_a,csr_ai,csr_aj
        end subroutine
        subroutine external_lu_solver(x) bind(c,name="external_lu_solver")
          import :: c_ptr
          type(c_ptr), value :: x
        end subroutine
        subroutine unload_external_lu_solver() &
                   bind(c,name="unload_external_lu_solver")
        end subroutine
      end interface
end module
