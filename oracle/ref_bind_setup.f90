! C-callable wrappers around the reference's set-up kernels,
! initialise_chunk_kernel (initialise_chunk_kernel.f90:26-182) and
! generate_chunk_kernel (generate_chunk_kernel.f90:30-211).
!
! TEST INFRASTRUCTURE ONLY; built with ref_bind.f90 into
! oracle/_ref/libclover_ref.so.  Scalars arrive by value, arrays as base
! addresses of Fortran column-major storage.

MODULE ref_bind_setup
  USE ISO_C_BINDING
  IMPLICIT NONE
CONTAINS

  SUBROUTINE ref_initialise_chunk_kernel(x_min,x_max,y_min,y_max,z_min,z_max,xmin,ymin,zmin,dx,dy,dz, &
      vertexx,vertexdx,vertexy,vertexdy,vertexz,vertexdz,cellx,celldx,celly,celldy,cellz,celldz,      &
      volume,xarea,yarea,zarea) BIND(C,NAME='ref_initialise_chunk_kernel')
    USE initialise_chunk_kernel_module, ONLY: initialise_chunk_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(C_DOUBLE), VALUE :: xmin,ymin,zmin,dx,dy,dz
    REAL(C_DOUBLE) :: vertexx(*),vertexdx(*),vertexy(*),vertexdy(*),vertexz(*),vertexdz(*)
    REAL(C_DOUBLE) :: cellx(*),celldx(*),celly(*),celldy(*),cellz(*),celldz(*)
    REAL(C_DOUBLE) :: volume(*),xarea(*),yarea(*),zarea(*)
    CALL initialise_chunk_kernel(x_min,x_max,y_min,y_max,z_min,z_max,xmin,ymin,zmin,dx,dy,dz, &
                                 vertexx,vertexdx,vertexy,vertexdy,vertexz,vertexdz,          &
                                 cellx,celldx,celly,celldy,cellz,celldz,volume,xarea,yarea,zarea)
  END SUBROUTINE ref_initialise_chunk_kernel

  ! g_rect, g_circ and g_point are data.f90:90-92's values, passed as
  ! generate_chunk.f90:83-85 passes them.
  SUBROUTINE ref_generate_chunk_kernel(x_min,x_max,y_min,y_max,z_min,z_max,vertexx,vertexy,vertexz, &
      cellx,celly,cellz,density0,energy0,xvel0,yvel0,zvel0,number_of_states,state_density,         &
      state_energy,state_xvel,state_yvel,state_zvel,state_xmin,state_xmax,state_ymin,state_ymax,   &
      state_zmin,state_zmax,state_radius,state_geometry) BIND(C,NAME='ref_generate_chunk_kernel')
    USE generate_chunk_kernel_module, ONLY: generate_chunk_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max,number_of_states
    REAL(C_DOUBLE) :: vertexx(*),vertexy(*),vertexz(*),cellx(*),celly(*),cellz(*)
    REAL(C_DOUBLE) :: density0(*),energy0(*),xvel0(*),yvel0(*),zvel0(*)
    REAL(C_DOUBLE) :: state_density(*),state_energy(*),state_xvel(*),state_yvel(*),state_zvel(*)
    REAL(C_DOUBLE) :: state_xmin(*),state_xmax(*),state_ymin(*),state_ymax(*),state_zmin(*),state_zmax(*)
    REAL(C_DOUBLE) :: state_radius(*)
    INTEGER(C_INT) :: state_geometry(*)
    CALL generate_chunk_kernel(x_min,x_max,y_min,y_max,z_min,z_max,vertexx,vertexy,vertexz, &
                               cellx,celly,cellz,density0,energy0,xvel0,yvel0,zvel0,          &
                               number_of_states,state_density,state_energy,state_xvel,        &
                               state_yvel,state_zvel,state_xmin,state_xmax,state_ymin,        &
                               state_ymax,state_zmin,state_zmax,state_radius,state_geometry,  &
                               1,2,3)
  END SUBROUTINE ref_generate_chunk_kernel

END MODULE ref_bind_setup
