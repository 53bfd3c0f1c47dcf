! Fortran side of the drop-in boundary.
!
! Each module below has the name of a reference kernel module, and it
! exports the subroutine the reference drivers call, with the same name and
! arguments (e.g. ideal_gas_kernel_module / ideal_gas_kernel,
! ideal_gas_kernel.f90:23-27).  Linking these modules in place of the
! reference's *_kernel.f90 files moves the hydro kernels onto the GPU
! without changing a line of the callers (hydro, timestep, PdV, advection,
! advec_cell_driver, advec_mom_driver, reset_field, revert, ideal_gas,
! viscosity, calc_dt, update_halo, update_tile_halo, field_summary).
!
! Each shim forwards to the C entry point of include/clover_hip.h:
!   integers and reals       by value
!   LOGICAL                   as an int
!   arrays                    base address of the caller's column-major
!                             array (assumed size, sequence association, so
!                             no copy is made)
!   outputs                   by reference
! A negative return code stops the program with the library's message.
! Call clover_hip_shutdown from clover_finalize (clover.f90:58-68) to free
! the device copies.

MODULE clover_hip_c
  USE ISO_C_BINDING
  IMPLICIT NONE

  INTERFACE
    FUNCTION c_last_error() BIND(C, NAME='clover_hip_last_error')
      IMPORT :: C_PTR
      TYPE(C_PTR) :: c_last_error
    END FUNCTION c_last_error

    INTEGER(C_INT) FUNCTION c_finalize() BIND(C, NAME='clover_hip_finalize')
      IMPORT :: C_INT
    END FUNCTION c_finalize

    INTEGER(C_INT) FUNCTION c_fetch(host) BIND(C, NAME='clover_hip_fetch')
      IMPORT :: C_INT, C_DOUBLE
      REAL(C_DOUBLE) :: host(*)
    END FUNCTION c_fetch

    INTEGER(C_INT) FUNCTION c_ideal_gas(x_min,x_max,y_min,y_max,z_min,z_max,density,energy,pressure,soundspeed) &
        BIND(C, NAME='clover_hip_ideal_gas_kernel')
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
      REAL(C_DOUBLE) :: density(*),energy(*),pressure(*),soundspeed(*)
    END FUNCTION c_ideal_gas

    INTEGER(C_INT) FUNCTION c_viscosity(x_min,x_max,y_min,y_max,z_min,z_max,xarea,yarea,zarea, &
        celldx,celldy,celldz,density0,pressure,viscosity,xvel0,yvel0,zvel0)                   &
        BIND(C, NAME='clover_hip_viscosity_kernel')
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
      REAL(C_DOUBLE) :: xarea(*),yarea(*),zarea(*),celldx(*),celldy(*),celldz(*)
      REAL(C_DOUBLE) :: density0(*),pressure(*),viscosity(*),xvel0(*),yvel0(*),zvel0(*)
    END FUNCTION c_viscosity

    INTEGER(C_INT) FUNCTION c_calc_dt(x_min,x_max,y_min,y_max,z_min,z_max,g_small,g_big,dtmin,           &
        dtc_safe,dtu_safe,dtv_safe,dtw_safe,dtdiv_safe,xarea,yarea,zarea,cellx,celly,cellz,              &
        celldx,celldy,celldz,volume,density0,energy0,pressure,viscosity_a,soundspeed,xvel0,yvel0,zvel0, &
        dt_min,dt_min_val,dtl_control,xl_pos,yl_pos,zl_pos,jldt,kldt,lldt,small)                        &
        BIND(C, NAME='clover_hip_calc_dt_kernel')
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
      REAL(C_DOUBLE), VALUE :: g_small,g_big,dtmin,dtc_safe,dtu_safe,dtv_safe,dtw_safe,dtdiv_safe
      REAL(C_DOUBLE) :: xarea(*),yarea(*),zarea(*),cellx(*),celly(*),cellz(*),celldx(*),celldy(*),celldz(*)
      REAL(C_DOUBLE) :: volume(*),density0(*),energy0(*),pressure(*),viscosity_a(*),soundspeed(*)
      REAL(C_DOUBLE) :: xvel0(*),yvel0(*),zvel0(*),dt_min(*)
      REAL(C_DOUBLE) :: dt_min_val,xl_pos,yl_pos,zl_pos
      INTEGER(C_INT) :: dtl_control,jldt,kldt,lldt,small
    END FUNCTION c_calc_dt

    INTEGER(C_INT) FUNCTION c_pdv(predict,x_min,x_max,y_min,y_max,z_min,z_max,dt,xarea,yarea,zarea,volume, &
        density0,density1,energy0,energy1,pressure,viscosity,xvel0,xvel1,yvel0,yvel1,zvel0,zvel1)          &
        BIND(C, NAME='clover_hip_PdV_kernel')
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: predict,x_min,x_max,y_min,y_max,z_min,z_max
      REAL(C_DOUBLE), VALUE :: dt
      REAL(C_DOUBLE) :: xarea(*),yarea(*),zarea(*),volume(*),density0(*),density1(*),energy0(*),energy1(*)
      REAL(C_DOUBLE) :: pressure(*),viscosity(*),xvel0(*),xvel1(*),yvel0(*),yvel1(*),zvel0(*),zvel1(*)
    END FUNCTION c_pdv

    INTEGER(C_INT) FUNCTION c_revert(x_min,x_max,y_min,y_max,z_min,z_max,density0,density1,energy0,energy1) &
        BIND(C, NAME='clover_hip_revert_kernel')
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
      REAL(C_DOUBLE) :: density0(*),density1(*),energy0(*),energy1(*)
    END FUNCTION c_revert

    INTEGER(C_INT) FUNCTION c_accelerate(x_min,x_max,y_min,y_max,z_min,z_max,dt,xarea,yarea,zarea,volume, &
        density0,pressure,viscosity,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1)                                 &
        BIND(C, NAME='clover_hip_accelerate_kernel')
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
      REAL(C_DOUBLE), VALUE :: dt
      REAL(C_DOUBLE) :: xarea(*),yarea(*),zarea(*),volume(*),density0(*),pressure(*),viscosity(*)
      REAL(C_DOUBLE) :: xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*)
    END FUNCTION c_accelerate

    INTEGER(C_INT) FUNCTION c_flux_calc(x_min,x_max,y_min,y_max,z_min,z_max,dt,xarea,yarea,zarea, &
        xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y,vol_flux_z)                    &
        BIND(C, NAME='clover_hip_flux_calc_kernel')
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
      REAL(C_DOUBLE), VALUE :: dt
      REAL(C_DOUBLE) :: xarea(*),yarea(*),zarea(*),xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*)
      REAL(C_DOUBLE) :: vol_flux_x(*),vol_flux_y(*),vol_flux_z(*)
    END FUNCTION c_flux_calc

    INTEGER(C_INT) FUNCTION c_advec_cell(x_min,x_max,y_min,y_max,z_min,z_max,advect_x,dir,sweep_number, &
        vertexdx,vertexdy,vertexdz,volume,density1,energy1,mass_flux_x,vol_flux_x,mass_flux_y,          &
        vol_flux_y,mass_flux_z,vol_flux_z,pre_vol,post_vol,pre_mass,post_mass,advec_vol,post_ener,     &
        ener_flux) BIND(C, NAME='clover_hip_advec_cell_kernel')
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max,advect_x,dir,sweep_number
      REAL(C_DOUBLE) :: vertexdx(*),vertexdy(*),vertexdz(*),volume(*),density1(*),energy1(*)
      REAL(C_DOUBLE) :: mass_flux_x(*),vol_flux_x(*),mass_flux_y(*),vol_flux_y(*),mass_flux_z(*),vol_flux_z(*)
      REAL(C_DOUBLE) :: pre_vol(*),post_vol(*),pre_mass(*),post_mass(*),advec_vol(*),post_ener(*),ener_flux(*)
    END FUNCTION c_advec_cell

    INTEGER(C_INT) FUNCTION c_advec_mom(x_min,x_max,y_min,y_max,z_min,z_max,vel1,mass_flux_x,vol_flux_x, &
        mass_flux_y,vol_flux_y,mass_flux_z,vol_flux_z,volume,density1,node_flux,node_mass_post,         &
        node_mass_pre,advec_vel,mom_flux,pre_vol,post_vol,celldx,celldy,celldz,advect_x,which_vel,      &
        sweep_number,direction) BIND(C, NAME='clover_hip_advec_mom_kernel')
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max,advect_x,which_vel,sweep_number,direction
      REAL(C_DOUBLE) :: vel1(*),mass_flux_x(*),vol_flux_x(*),mass_flux_y(*),vol_flux_y(*)
      REAL(C_DOUBLE) :: mass_flux_z(*),vol_flux_z(*),volume(*),density1(*),node_flux(*),node_mass_post(*)
      REAL(C_DOUBLE) :: node_mass_pre(*),advec_vel(*),mom_flux(*),pre_vol(*),post_vol(*)
      REAL(C_DOUBLE) :: celldx(*),celldy(*),celldz(*)
    END FUNCTION c_advec_mom

    INTEGER(C_INT) FUNCTION c_reset_field(x_min,x_max,y_min,y_max,z_min,z_max,density0,density1, &
        energy0,energy1,xvel0,xvel1,yvel0,yvel1,zvel0,zvel1) BIND(C, NAME='clover_hip_reset_field_kernel')
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
      REAL(C_DOUBLE) :: density0(*),density1(*),energy0(*),energy1(*)
      REAL(C_DOUBLE) :: xvel0(*),xvel1(*),yvel0(*),yvel1(*),zvel0(*),zvel1(*)
    END FUNCTION c_reset_field

    INTEGER(C_INT) FUNCTION c_update_halo(x_min,x_max,y_min,y_max,z_min,z_max,chunk_neighbours,         &
        tile_neighbours,density0,energy0,pressure,viscosity,soundspeed,density1,energy1,xvel0,yvel0,    &
        zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y,vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,   &
        fields,depth) BIND(C, NAME='clover_hip_update_halo_kernel')
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max,depth
      INTEGER(C_INT) :: chunk_neighbours(*),tile_neighbours(*),fields(*)
      REAL(C_DOUBLE) :: density0(*),energy0(*),pressure(*),viscosity(*),soundspeed(*),density1(*),energy1(*)
      REAL(C_DOUBLE) :: xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*)
      REAL(C_DOUBLE) :: vol_flux_x(*),vol_flux_y(*),vol_flux_z(*),mass_flux_x(*),mass_flux_y(*),mass_flux_z(*)
    END FUNCTION c_update_halo

    INTEGER(C_INT) FUNCTION c_field_summary(x_min,x_max,y_min,y_max,z_min,z_max,volume,density0,energy0, &
        pressure,xvel0,yvel0,zvel0,vol,mass,ie,ke,press) BIND(C, NAME='clover_hip_field_summary_kernel')
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
      REAL(C_DOUBLE) :: volume(*),density0(*),energy0(*),pressure(*),xvel0(*),yvel0(*),zvel0(*)
      REAL(C_DOUBLE) :: vol,mass,ie,ke,press
    END FUNCTION c_field_summary
  END INTERFACE

  ! The six tile-halo entry points share one argument list
  ! (update_tile_halo_kernel.f90:31-75); they differ only in the face.
  ABSTRACT INTERFACE
    INTEGER(C_INT) FUNCTION c_tile_face(x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0,pressure,  &
        viscosity,soundspeed,density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y, &
        vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,n_xmin,n_xmax,n_ymin,n_ymax,n_zmin,n_zmax,        &
        n_density0,n_energy0,n_pressure,n_viscosity,n_soundspeed,n_density1,n_energy1,n_xvel0,n_yvel0,   &
        n_zvel0,n_xvel1,n_yvel1,n_zvel1,n_vol_flux_x,n_vol_flux_y,n_vol_flux_z,n_mass_flux_x,            &
        n_mass_flux_y,n_mass_flux_z,fields,depth) BIND(C)
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max,n_xmin,n_xmax,n_ymin,n_ymax,n_zmin,n_zmax
      INTEGER(C_INT), VALUE :: depth
      INTEGER(C_INT) :: fields(*)
      REAL(C_DOUBLE) :: density0(*),energy0(*),pressure(*),viscosity(*),soundspeed(*),density1(*),energy1(*)
      REAL(C_DOUBLE) :: xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*)
      REAL(C_DOUBLE) :: vol_flux_x(*),vol_flux_y(*),vol_flux_z(*),mass_flux_x(*),mass_flux_y(*),mass_flux_z(*)
      REAL(C_DOUBLE) :: n_density0(*),n_energy0(*),n_pressure(*),n_viscosity(*),n_soundspeed(*)
      REAL(C_DOUBLE) :: n_density1(*),n_energy1(*),n_xvel0(*),n_yvel0(*),n_zvel0(*),n_xvel1(*),n_yvel1(*)
      REAL(C_DOUBLE) :: n_zvel1(*),n_vol_flux_x(*),n_vol_flux_y(*),n_vol_flux_z(*)
      REAL(C_DOUBLE) :: n_mass_flux_x(*),n_mass_flux_y(*),n_mass_flux_z(*)
    END FUNCTION c_tile_face
  END INTERFACE

  PROCEDURE(c_tile_face), BIND(C, NAME='clover_hip_update_tile_halo_left_kernel') :: c_tile_left
  PROCEDURE(c_tile_face), BIND(C, NAME='clover_hip_update_tile_halo_right_kernel') :: c_tile_right
  PROCEDURE(c_tile_face), BIND(C, NAME='clover_hip_update_tile_halo_top_kernel') :: c_tile_top
  PROCEDURE(c_tile_face), BIND(C, NAME='clover_hip_update_tile_halo_bottom_kernel') :: c_tile_bottom
  PROCEDURE(c_tile_face), BIND(C, NAME='clover_hip_update_tile_halo_front_kernel') :: c_tile_front
  PROCEDURE(c_tile_face), BIND(C, NAME='clover_hip_update_tile_halo_back_kernel') :: c_tile_back

CONTAINS

  ! Stops the program with the library's message when rc is negative
  ! (the reference's report_error path, report.f90:29-56, aborts likewise).
  SUBROUTINE clover_hip_check(rc, what)
    INTEGER(C_INT), INTENT(IN) :: rc
    CHARACTER(LEN=*), INTENT(IN) :: what
    CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
    INTEGER :: n
    IF (rc >= 0) RETURN
    CALL C_F_POINTER(c_last_error(), msg, [4096])
    n = 0
    DO WHILE (n < 4096)
      IF (msg(n + 1) == C_NULL_CHAR) EXIT
      n = n + 1
    END DO
    WRITE(*, '(A, A, A, I0, A, 4096A1)') 'clover_hip: ', what, ' failed (', rc, '): ', msg(1:n)
    ERROR STOP 1
  END SUBROUTINE clover_hip_check

  ! Frees every device copy; for clover_finalize.
  SUBROUTINE clover_hip_shutdown()
    CALL clover_hip_check(c_finalize(), 'clover_hip_finalize')
  END SUBROUTINE clover_hip_shutdown

  ! Copies the device copy of a field back into its host array (for visit
  ! output or any other host reader).
  SUBROUTINE clover_hip_fetch_field(host)
    REAL(C_DOUBLE) :: host(*)
    CALL clover_hip_check(c_fetch(host), 'clover_hip_fetch')
  END SUBROUTINE clover_hip_fetch_field

  INTEGER(C_INT) FUNCTION as_int(flag)
    LOGICAL, INTENT(IN) :: flag
    as_int = MERGE(1_C_INT, 0_C_INT, flag)
  END FUNCTION as_int

END MODULE clover_hip_c


MODULE ideal_gas_kernel_module
  USE clover_hip_c
  IMPLICIT NONE
CONTAINS
  SUBROUTINE ideal_gas_kernel(x_min,x_max,y_min,y_max,z_min,z_max,density,energy,pressure,soundspeed)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(KIND=8) :: density(*),energy(*),pressure(*),soundspeed(*)
    CALL clover_hip_check(c_ideal_gas(x_min,x_max,y_min,y_max,z_min,z_max,density,energy,pressure,soundspeed), &
                          'ideal_gas_kernel')
  END SUBROUTINE ideal_gas_kernel
END MODULE ideal_gas_kernel_module


MODULE viscosity_kernel_module
  USE clover_hip_c
  IMPLICIT NONE
CONTAINS
  SUBROUTINE viscosity_kernel(x_min,x_max,y_min,y_max,z_min,z_max,xarea,yarea,zarea, &
                              celldx,celldy,celldz,density0,pressure,viscosity,xvel0,yvel0,zvel0)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(KIND=8) :: xarea(*),yarea(*),zarea(*),celldx(*),celldy(*),celldz(*)
    REAL(KIND=8) :: density0(*),pressure(*),viscosity(*),xvel0(*),yvel0(*),zvel0(*)
    CALL clover_hip_check(c_viscosity(x_min,x_max,y_min,y_max,z_min,z_max,xarea,yarea,zarea,celldx,celldy, &
                                      celldz,density0,pressure,viscosity,xvel0,yvel0,zvel0), 'viscosity_kernel')
  END SUBROUTINE viscosity_kernel
END MODULE viscosity_kernel_module


MODULE calc_dt_kernel_module
  USE clover_hip_c
  IMPLICIT NONE
CONTAINS
  SUBROUTINE calc_dt_kernel(x_min,x_max,y_min,y_max,z_min,z_max,g_small,g_big,dtmin,          &
                            dtc_safe,dtu_safe,dtv_safe,dtw_safe,dtdiv_safe,xarea,yarea,zarea, &
                            cellx,celly,cellz,celldx,celldy,celldz,volume,density0,energy0,   &
                            pressure,viscosity_a,soundspeed,xvel0,yvel0,zvel0,dt_min,         &
                            dt_min_val,dtl_control,xl_pos,yl_pos,zl_pos,jldt,kldt,lldt,small)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(KIND=8) :: g_small,g_big,dtmin,dtc_safe,dtu_safe,dtv_safe,dtw_safe,dtdiv_safe
    REAL(KIND=8) :: xarea(*),yarea(*),zarea(*),cellx(*),celly(*),cellz(*),celldx(*),celldy(*),celldz(*)
    REAL(KIND=8) :: volume(*),density0(*),energy0(*),pressure(*),viscosity_a(*),soundspeed(*)
    REAL(KIND=8) :: xvel0(*),yvel0(*),zvel0(*),dt_min(*)
    REAL(KIND=8) :: dt_min_val,xl_pos,yl_pos,zl_pos
    INTEGER :: dtl_control,jldt,kldt,lldt,small
    CALL clover_hip_check(c_calc_dt(x_min,x_max,y_min,y_max,z_min,z_max,g_small,g_big,dtmin,dtc_safe,    &
                                    dtu_safe,dtv_safe,dtw_safe,dtdiv_safe,xarea,yarea,zarea,cellx,celly, &
                                    cellz,celldx,celldy,celldz,volume,density0,energy0,pressure,         &
                                    viscosity_a,soundspeed,xvel0,yvel0,zvel0,dt_min,dt_min_val,          &
                                    dtl_control,xl_pos,yl_pos,zl_pos,jldt,kldt,lldt,small), 'calc_dt_kernel')
  END SUBROUTINE calc_dt_kernel
END MODULE calc_dt_kernel_module


MODULE PdV_kernel_module
  USE clover_hip_c
  IMPLICIT NONE
CONTAINS
  SUBROUTINE PdV_kernel(predict,x_min,x_max,y_min,y_max,z_min,z_max,dt,xarea,yarea,zarea,volume, &
                        density0,density1,energy0,energy1,pressure,viscosity,                    &
                        xvel0,xvel1,yvel0,yvel1,zvel0,zvel1)
    LOGICAL :: predict
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(KIND=8) :: dt
    REAL(KIND=8) :: xarea(*),yarea(*),zarea(*),volume(*),density0(*),density1(*),energy0(*),energy1(*)
    REAL(KIND=8) :: pressure(*),viscosity(*),xvel0(*),xvel1(*),yvel0(*),yvel1(*),zvel0(*),zvel1(*)
    CALL clover_hip_check(c_pdv(as_int(predict),x_min,x_max,y_min,y_max,z_min,z_max,dt,xarea,yarea,zarea, &
                                volume,density0,density1,energy0,energy1,pressure,viscosity,xvel0,xvel1,   &
                                yvel0,yvel1,zvel0,zvel1), 'PdV_kernel')
  END SUBROUTINE PdV_kernel
END MODULE PdV_kernel_module


MODULE revert_kernel_module
  USE clover_hip_c
  IMPLICIT NONE
CONTAINS
  SUBROUTINE revert_kernel(x_min,x_max,y_min,y_max,z_min,z_max,density0,density1,energy0,energy1)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(KIND=8) :: density0(*),density1(*),energy0(*),energy1(*)
    CALL clover_hip_check(c_revert(x_min,x_max,y_min,y_max,z_min,z_max,density0,density1,energy0,energy1), &
                          'revert_kernel')
  END SUBROUTINE revert_kernel
END MODULE revert_kernel_module


MODULE accelerate_kernel_module
  USE clover_hip_c
  IMPLICIT NONE
CONTAINS
  SUBROUTINE accelerate_kernel(x_min,x_max,y_min,y_max,z_min,z_max,dt,xarea,yarea,zarea,volume, &
                               density0,pressure,viscosity,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(KIND=8) :: dt
    REAL(KIND=8) :: xarea(*),yarea(*),zarea(*),volume(*),density0(*),pressure(*),viscosity(*)
    REAL(KIND=8) :: xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*)
    CALL clover_hip_check(c_accelerate(x_min,x_max,y_min,y_max,z_min,z_max,dt,xarea,yarea,zarea,volume,    &
                                       density0,pressure,viscosity,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1), &
                          'accelerate_kernel')
  END SUBROUTINE accelerate_kernel
END MODULE accelerate_kernel_module


MODULE flux_calc_kernel_module
  USE clover_hip_c
  IMPLICIT NONE
CONTAINS
  SUBROUTINE flux_calc_kernel(x_min,x_max,y_min,y_max,z_min,z_max,dt,xarea,yarea,zarea, &
                              xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y,vol_flux_z)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(KIND=8) :: dt
    REAL(KIND=8) :: xarea(*),yarea(*),zarea(*),xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*)
    REAL(KIND=8) :: vol_flux_x(*),vol_flux_y(*),vol_flux_z(*)
    CALL clover_hip_check(c_flux_calc(x_min,x_max,y_min,y_max,z_min,z_max,dt,xarea,yarea,zarea,xvel0,yvel0, &
                                      zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y,vol_flux_z),            &
                          'flux_calc_kernel')
  END SUBROUTINE flux_calc_kernel
END MODULE flux_calc_kernel_module


MODULE advec_cell_kernel_module
  USE clover_hip_c
  IMPLICIT NONE
CONTAINS
  SUBROUTINE advec_cell_kernel(x_min,x_max,y_min,y_max,z_min,z_max,advect_x,dir,sweep_number,        &
                               vertexdx,vertexdy,vertexdz,volume,density1,energy1,mass_flux_x,       &
                               vol_flux_x,mass_flux_y,vol_flux_y,mass_flux_z,vol_flux_z,pre_vol,     &
                               post_vol,pre_mass,post_mass,advec_vol,post_ener,ener_flux)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max,dir,sweep_number
    LOGICAL :: advect_x
    REAL(KIND=8) :: vertexdx(*),vertexdy(*),vertexdz(*),volume(*),density1(*),energy1(*)
    REAL(KIND=8) :: mass_flux_x(*),vol_flux_x(*),mass_flux_y(*),vol_flux_y(*),mass_flux_z(*),vol_flux_z(*)
    REAL(KIND=8) :: pre_vol(*),post_vol(*),pre_mass(*),post_mass(*),advec_vol(*),post_ener(*),ener_flux(*)
    CALL clover_hip_check(c_advec_cell(x_min,x_max,y_min,y_max,z_min,z_max,as_int(advect_x),dir,          &
                                       sweep_number,vertexdx,vertexdy,vertexdz,volume,density1,energy1,   &
                                       mass_flux_x,vol_flux_x,mass_flux_y,vol_flux_y,mass_flux_z,         &
                                       vol_flux_z,pre_vol,post_vol,pre_mass,post_mass,advec_vol,          &
                                       post_ener,ener_flux), 'advec_cell_kernel')
  END SUBROUTINE advec_cell_kernel
END MODULE advec_cell_kernel_module


MODULE advec_mom_kernel_mod
  USE clover_hip_c
  IMPLICIT NONE
CONTAINS
  SUBROUTINE advec_mom_kernel(x_min,x_max,y_min,y_max,z_min,z_max,vel1,mass_flux_x,vol_flux_x,         &
                              mass_flux_y,vol_flux_y,mass_flux_z,vol_flux_z,volume,density1,node_flux, &
                              node_mass_post,node_mass_pre,advec_vel,mom_flux,pre_vol,post_vol,        &
                              celldx,celldy,celldz,advect_x,which_vel,sweep_number,direction)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max,which_vel,sweep_number,direction
    LOGICAL :: advect_x
    REAL(KIND=8) :: vel1(*),mass_flux_x(*),vol_flux_x(*),mass_flux_y(*),vol_flux_y(*)
    REAL(KIND=8) :: mass_flux_z(*),vol_flux_z(*),volume(*),density1(*),node_flux(*),node_mass_post(*)
    REAL(KIND=8) :: node_mass_pre(*),advec_vel(*),mom_flux(*),pre_vol(*),post_vol(*)
    REAL(KIND=8) :: celldx(*),celldy(*),celldz(*)
    CALL clover_hip_check(c_advec_mom(x_min,x_max,y_min,y_max,z_min,z_max,vel1,mass_flux_x,vol_flux_x,     &
                                      mass_flux_y,vol_flux_y,mass_flux_z,vol_flux_z,volume,density1,       &
                                      node_flux,node_mass_post,node_mass_pre,advec_vel,mom_flux,pre_vol,   &
                                      post_vol,celldx,celldy,celldz,as_int(advect_x),which_vel,            &
                                      sweep_number,direction), 'advec_mom_kernel')
  END SUBROUTINE advec_mom_kernel
END MODULE advec_mom_kernel_mod


MODULE reset_field_kernel_module
  USE clover_hip_c
  IMPLICIT NONE
CONTAINS
  SUBROUTINE reset_field_kernel(x_min,x_max,y_min,y_max,z_min,z_max,density0,density1,energy0,energy1, &
                                xvel0,xvel1,yvel0,yvel1,zvel0,zvel1)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(KIND=8) :: density0(*),density1(*),energy0(*),energy1(*)
    REAL(KIND=8) :: xvel0(*),xvel1(*),yvel0(*),yvel1(*),zvel0(*),zvel1(*)
    CALL clover_hip_check(c_reset_field(x_min,x_max,y_min,y_max,z_min,z_max,density0,density1,energy0,  &
                                        energy1,xvel0,xvel1,yvel0,yvel1,zvel0,zvel1), 'reset_field_kernel')
  END SUBROUTINE reset_field_kernel
END MODULE reset_field_kernel_module


MODULE update_halo_kernel_module
  USE clover_hip_c
  IMPLICIT NONE
CONTAINS
  SUBROUTINE update_halo_kernel(x_min,x_max,y_min,y_max,z_min,z_max,chunk_neighbours,tile_neighbours,   &
                                density0,energy0,pressure,viscosity,soundspeed,density1,energy1,xvel0,  &
                                yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y,vol_flux_z,         &
                                mass_flux_x,mass_flux_y,mass_flux_z,fields,depth)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max,depth
    INTEGER :: chunk_neighbours(*),tile_neighbours(*),fields(*)
    REAL(KIND=8) :: density0(*),energy0(*),pressure(*),viscosity(*),soundspeed(*),density1(*),energy1(*)
    REAL(KIND=8) :: xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*)
    REAL(KIND=8) :: vol_flux_x(*),vol_flux_y(*),vol_flux_z(*),mass_flux_x(*),mass_flux_y(*),mass_flux_z(*)
    CALL clover_hip_check(c_update_halo(x_min,x_max,y_min,y_max,z_min,z_max,chunk_neighbours,               &
                                        tile_neighbours,density0,energy0,pressure,viscosity,soundspeed,     &
                                        density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,    &
                                        vol_flux_y,vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,fields,   &
                                        depth), 'update_halo_kernel')
  END SUBROUTINE update_halo_kernel
END MODULE update_halo_kernel_module


MODULE update_tile_halo_kernel_module
  USE clover_hip_c
  IMPLICIT NONE
CONTAINS
  ! The six face subroutines (update_tile_halo_kernel.f90:31, 361, 690,
  ! 1022, 1351, 1674) share one argument list; each forwards to its face.

  SUBROUTINE update_tile_halo_left_kernel(x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0, &
      pressure,viscosity,soundspeed,density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x, &
      vol_flux_y,vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,left_xmin,left_xmax,left_ymin, &
      left_ymax,left_zmin,left_zmax,left_density0,left_energy0,left_pressure,left_viscosity, &
      left_soundspeed,left_density1,left_energy1,left_xvel0,left_yvel0,left_zvel0,left_xvel1, &
      left_yvel1,left_zvel1,left_vol_flux_x,left_vol_flux_y,left_vol_flux_z,left_mass_flux_x, &
      left_mass_flux_y,left_mass_flux_z,fields,depth)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max,depth,fields(*)
    INTEGER :: left_xmin,left_xmax,left_ymin,left_ymax,left_zmin,left_zmax
    REAL(KIND=8) :: density0(*),energy0(*),pressure(*),viscosity(*),soundspeed(*),density1(*), &
      energy1(*),xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*),vol_flux_x(*),vol_flux_y(*), &
      vol_flux_z(*),mass_flux_x(*),mass_flux_y(*),mass_flux_z(*)
    REAL(KIND=8) :: left_density0(*),left_energy0(*),left_pressure(*),left_viscosity(*), &
      left_soundspeed(*),left_density1(*),left_energy1(*),left_xvel0(*),left_yvel0(*),left_zvel0(*), &
      left_xvel1(*),left_yvel1(*),left_zvel1(*),left_vol_flux_x(*),left_vol_flux_y(*), &
      left_vol_flux_z(*),left_mass_flux_x(*),left_mass_flux_y(*),left_mass_flux_z(*)
    CALL clover_hip_check(c_tile_left(x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0,pressure, &
      viscosity,soundspeed,density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x, &
      vol_flux_y,vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,left_xmin,left_xmax,left_ymin, &
      left_ymax,left_zmin,left_zmax,left_density0,left_energy0,left_pressure,left_viscosity, &
      left_soundspeed,left_density1,left_energy1,left_xvel0,left_yvel0,left_zvel0,left_xvel1, &
      left_yvel1,left_zvel1,left_vol_flux_x,left_vol_flux_y,left_vol_flux_z,left_mass_flux_x, &
      left_mass_flux_y,left_mass_flux_z,fields,depth), &
                          'update_tile_halo_left_kernel')
  END SUBROUTINE update_tile_halo_left_kernel

  SUBROUTINE update_tile_halo_right_kernel(x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0, &
      pressure,viscosity,soundspeed,density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x, &
      vol_flux_y,vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,right_xmin,right_xmax,right_ymin, &
      right_ymax,right_zmin,right_zmax,right_density0,right_energy0,right_pressure,right_viscosity, &
      right_soundspeed,right_density1,right_energy1,right_xvel0,right_yvel0,right_zvel0,right_xvel1, &
      right_yvel1,right_zvel1,right_vol_flux_x,right_vol_flux_y,right_vol_flux_z,right_mass_flux_x, &
      right_mass_flux_y,right_mass_flux_z,fields,depth)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max,depth,fields(*)
    INTEGER :: right_xmin,right_xmax,right_ymin,right_ymax,right_zmin,right_zmax
    REAL(KIND=8) :: density0(*),energy0(*),pressure(*),viscosity(*),soundspeed(*),density1(*), &
      energy1(*),xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*),vol_flux_x(*),vol_flux_y(*), &
      vol_flux_z(*),mass_flux_x(*),mass_flux_y(*),mass_flux_z(*)
    REAL(KIND=8) :: right_density0(*),right_energy0(*),right_pressure(*),right_viscosity(*), &
      right_soundspeed(*),right_density1(*),right_energy1(*),right_xvel0(*),right_yvel0(*), &
      right_zvel0(*),right_xvel1(*),right_yvel1(*),right_zvel1(*),right_vol_flux_x(*), &
      right_vol_flux_y(*),right_vol_flux_z(*),right_mass_flux_x(*),right_mass_flux_y(*), &
      right_mass_flux_z(*)
    CALL clover_hip_check(c_tile_right(x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0, &
      pressure,viscosity,soundspeed,density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x, &
      vol_flux_y,vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,right_xmin,right_xmax,right_ymin, &
      right_ymax,right_zmin,right_zmax,right_density0,right_energy0,right_pressure,right_viscosity, &
      right_soundspeed,right_density1,right_energy1,right_xvel0,right_yvel0,right_zvel0,right_xvel1, &
      right_yvel1,right_zvel1,right_vol_flux_x,right_vol_flux_y,right_vol_flux_z,right_mass_flux_x, &
      right_mass_flux_y,right_mass_flux_z,fields,depth), &
                          'update_tile_halo_right_kernel')
  END SUBROUTINE update_tile_halo_right_kernel

  SUBROUTINE update_tile_halo_top_kernel(x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0, &
      pressure,viscosity,soundspeed,density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x, &
      vol_flux_y,vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,top_xmin,top_xmax,top_ymin,top_ymax, &
      top_zmin,top_zmax,top_density0,top_energy0,top_pressure,top_viscosity,top_soundspeed, &
      top_density1,top_energy1,top_xvel0,top_yvel0,top_zvel0,top_xvel1,top_yvel1,top_zvel1, &
      top_vol_flux_x,top_vol_flux_y,top_vol_flux_z,top_mass_flux_x,top_mass_flux_y,top_mass_flux_z, &
      fields,depth)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max,depth,fields(*)
    INTEGER :: top_xmin,top_xmax,top_ymin,top_ymax,top_zmin,top_zmax
    REAL(KIND=8) :: density0(*),energy0(*),pressure(*),viscosity(*),soundspeed(*),density1(*), &
      energy1(*),xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*),vol_flux_x(*),vol_flux_y(*), &
      vol_flux_z(*),mass_flux_x(*),mass_flux_y(*),mass_flux_z(*)
    REAL(KIND=8) :: top_density0(*),top_energy0(*),top_pressure(*),top_viscosity(*), &
      top_soundspeed(*),top_density1(*),top_energy1(*),top_xvel0(*),top_yvel0(*),top_zvel0(*), &
      top_xvel1(*),top_yvel1(*),top_zvel1(*),top_vol_flux_x(*),top_vol_flux_y(*),top_vol_flux_z(*), &
      top_mass_flux_x(*),top_mass_flux_y(*),top_mass_flux_z(*)
    CALL clover_hip_check(c_tile_top(x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0,pressure, &
      viscosity,soundspeed,density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x, &
      vol_flux_y,vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,top_xmin,top_xmax,top_ymin,top_ymax, &
      top_zmin,top_zmax,top_density0,top_energy0,top_pressure,top_viscosity,top_soundspeed, &
      top_density1,top_energy1,top_xvel0,top_yvel0,top_zvel0,top_xvel1,top_yvel1,top_zvel1, &
      top_vol_flux_x,top_vol_flux_y,top_vol_flux_z,top_mass_flux_x,top_mass_flux_y,top_mass_flux_z, &
      fields,depth), &
                          'update_tile_halo_top_kernel')
  END SUBROUTINE update_tile_halo_top_kernel

  SUBROUTINE update_tile_halo_bottom_kernel(x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0, &
      pressure,viscosity,soundspeed,density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x, &
      vol_flux_y,vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,bottom_xmin,bottom_xmax,bottom_ymin, &
      bottom_ymax,bottom_zmin,bottom_zmax,bottom_density0,bottom_energy0,bottom_pressure, &
      bottom_viscosity,bottom_soundspeed,bottom_density1,bottom_energy1,bottom_xvel0,bottom_yvel0, &
      bottom_zvel0,bottom_xvel1,bottom_yvel1,bottom_zvel1,bottom_vol_flux_x,bottom_vol_flux_y, &
      bottom_vol_flux_z,bottom_mass_flux_x,bottom_mass_flux_y,bottom_mass_flux_z,fields,depth)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max,depth,fields(*)
    INTEGER :: bottom_xmin,bottom_xmax,bottom_ymin,bottom_ymax,bottom_zmin,bottom_zmax
    REAL(KIND=8) :: density0(*),energy0(*),pressure(*),viscosity(*),soundspeed(*),density1(*), &
      energy1(*),xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*),vol_flux_x(*),vol_flux_y(*), &
      vol_flux_z(*),mass_flux_x(*),mass_flux_y(*),mass_flux_z(*)
    REAL(KIND=8) :: bottom_density0(*),bottom_energy0(*),bottom_pressure(*),bottom_viscosity(*), &
      bottom_soundspeed(*),bottom_density1(*),bottom_energy1(*),bottom_xvel0(*),bottom_yvel0(*), &
      bottom_zvel0(*),bottom_xvel1(*),bottom_yvel1(*),bottom_zvel1(*),bottom_vol_flux_x(*), &
      bottom_vol_flux_y(*),bottom_vol_flux_z(*),bottom_mass_flux_x(*),bottom_mass_flux_y(*), &
      bottom_mass_flux_z(*)
    CALL clover_hip_check(c_tile_bottom(x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0, &
      pressure,viscosity,soundspeed,density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x, &
      vol_flux_y,vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,bottom_xmin,bottom_xmax,bottom_ymin, &
      bottom_ymax,bottom_zmin,bottom_zmax,bottom_density0,bottom_energy0,bottom_pressure, &
      bottom_viscosity,bottom_soundspeed,bottom_density1,bottom_energy1,bottom_xvel0,bottom_yvel0, &
      bottom_zvel0,bottom_xvel1,bottom_yvel1,bottom_zvel1,bottom_vol_flux_x,bottom_vol_flux_y, &
      bottom_vol_flux_z,bottom_mass_flux_x,bottom_mass_flux_y,bottom_mass_flux_z,fields,depth), &
                          'update_tile_halo_bottom_kernel')
  END SUBROUTINE update_tile_halo_bottom_kernel

  SUBROUTINE update_tile_halo_front_kernel(x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0, &
      pressure,viscosity,soundspeed,density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x, &
      vol_flux_y,vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,front_xmin,front_xmax,front_ymin, &
      front_ymax,front_zmin,front_zmax,front_density0,front_energy0,front_pressure,front_viscosity, &
      front_soundspeed,front_density1,front_energy1,front_xvel0,front_yvel0,front_zvel0,front_xvel1, &
      front_yvel1,front_zvel1,front_vol_flux_x,front_vol_flux_y,front_vol_flux_z,front_mass_flux_x, &
      front_mass_flux_y,front_mass_flux_z,fields,depth)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max,depth,fields(*)
    INTEGER :: front_xmin,front_xmax,front_ymin,front_ymax,front_zmin,front_zmax
    REAL(KIND=8) :: density0(*),energy0(*),pressure(*),viscosity(*),soundspeed(*),density1(*), &
      energy1(*),xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*),vol_flux_x(*),vol_flux_y(*), &
      vol_flux_z(*),mass_flux_x(*),mass_flux_y(*),mass_flux_z(*)
    REAL(KIND=8) :: front_density0(*),front_energy0(*),front_pressure(*),front_viscosity(*), &
      front_soundspeed(*),front_density1(*),front_energy1(*),front_xvel0(*),front_yvel0(*), &
      front_zvel0(*),front_xvel1(*),front_yvel1(*),front_zvel1(*),front_vol_flux_x(*), &
      front_vol_flux_y(*),front_vol_flux_z(*),front_mass_flux_x(*),front_mass_flux_y(*), &
      front_mass_flux_z(*)
    CALL clover_hip_check(c_tile_front(x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0, &
      pressure,viscosity,soundspeed,density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x, &
      vol_flux_y,vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,front_xmin,front_xmax,front_ymin, &
      front_ymax,front_zmin,front_zmax,front_density0,front_energy0,front_pressure,front_viscosity, &
      front_soundspeed,front_density1,front_energy1,front_xvel0,front_yvel0,front_zvel0,front_xvel1, &
      front_yvel1,front_zvel1,front_vol_flux_x,front_vol_flux_y,front_vol_flux_z,front_mass_flux_x, &
      front_mass_flux_y,front_mass_flux_z,fields,depth), &
                          'update_tile_halo_front_kernel')
  END SUBROUTINE update_tile_halo_front_kernel

  SUBROUTINE update_tile_halo_back_kernel(x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0, &
      pressure,viscosity,soundspeed,density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x, &
      vol_flux_y,vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,back_xmin,back_xmax,back_ymin, &
      back_ymax,back_zmin,back_zmax,back_density0,back_energy0,back_pressure,back_viscosity, &
      back_soundspeed,back_density1,back_energy1,back_xvel0,back_yvel0,back_zvel0,back_xvel1, &
      back_yvel1,back_zvel1,back_vol_flux_x,back_vol_flux_y,back_vol_flux_z,back_mass_flux_x, &
      back_mass_flux_y,back_mass_flux_z,fields,depth)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max,depth,fields(*)
    INTEGER :: back_xmin,back_xmax,back_ymin,back_ymax,back_zmin,back_zmax
    REAL(KIND=8) :: density0(*),energy0(*),pressure(*),viscosity(*),soundspeed(*),density1(*), &
      energy1(*),xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*),vol_flux_x(*),vol_flux_y(*), &
      vol_flux_z(*),mass_flux_x(*),mass_flux_y(*),mass_flux_z(*)
    REAL(KIND=8) :: back_density0(*),back_energy0(*),back_pressure(*),back_viscosity(*), &
      back_soundspeed(*),back_density1(*),back_energy1(*),back_xvel0(*),back_yvel0(*),back_zvel0(*), &
      back_xvel1(*),back_yvel1(*),back_zvel1(*),back_vol_flux_x(*),back_vol_flux_y(*), &
      back_vol_flux_z(*),back_mass_flux_x(*),back_mass_flux_y(*),back_mass_flux_z(*)
    CALL clover_hip_check(c_tile_back(x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0,pressure, &
      viscosity,soundspeed,density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x, &
      vol_flux_y,vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,back_xmin,back_xmax,back_ymin, &
      back_ymax,back_zmin,back_zmax,back_density0,back_energy0,back_pressure,back_viscosity, &
      back_soundspeed,back_density1,back_energy1,back_xvel0,back_yvel0,back_zvel0,back_xvel1, &
      back_yvel1,back_zvel1,back_vol_flux_x,back_vol_flux_y,back_vol_flux_z,back_mass_flux_x, &
      back_mass_flux_y,back_mass_flux_z,fields,depth), &
                          'update_tile_halo_back_kernel')
  END SUBROUTINE update_tile_halo_back_kernel
END MODULE update_tile_halo_kernel_module


MODULE field_summary_kernel_module
  USE clover_hip_c
  IMPLICIT NONE
CONTAINS
  SUBROUTINE field_summary_kernel(x_min,x_max,y_min,y_max,z_min,z_max,volume,density0,energy0,pressure, &
                                  xvel0,yvel0,zvel0,vol,mass,ie,ke,press)
    INTEGER :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(KIND=8) :: volume(*),density0(*),energy0(*),pressure(*),xvel0(*),yvel0(*),zvel0(*)
    REAL(KIND=8) :: vol,mass,ie,ke,press
    CALL clover_hip_check(c_field_summary(x_min,x_max,y_min,y_max,z_min,z_max,volume,density0,energy0,   &
                                          pressure,xvel0,yvel0,zvel0,vol,mass,ie,ke,press),              &
                          'field_summary_kernel')
  END SUBROUTINE field_summary_kernel
END MODULE field_summary_kernel_module
