! C-callable wrappers around the reference's six update_tile_halo_*_kernel
! subroutines (update_tile_halo_kernel.f90:31-1993).
!
! TEST INFRASTRUCTURE ONLY; built with ref_bind.f90 into
! oracle/_ref/libclover_ref.so.  Each wrapper takes the tile's bounds and 19
! field arrays, the neighbour's bounds and 19 arrays (same order), the 19
! field flags and the depth.

MODULE ref_bind_tile
  USE ISO_C_BINDING
  IMPLICIT NONE
CONTAINS

  SUBROUTINE ref_update_tile_halo_left_kernel( &
      x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0,pressure,viscosity,soundspeed, &
      density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y, &
      vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,n_xmin,n_xmax,n_ymin,n_ymax,n_zmin, &
      n_zmax,n_density0,n_energy0,n_pressure,n_viscosity,n_soundspeed,n_density1, &
      n_energy1,n_xvel0,n_yvel0,n_zvel0,n_xvel1,n_yvel1,n_zvel1,n_vol_flux_x,n_vol_flux_y, &
      n_vol_flux_z,n_mass_flux_x,n_mass_flux_y,n_mass_flux_z,fields,depth) BIND(C,NAME='ref_update_tile_halo_left_kernel')
    USE update_tile_halo_kernel_module, ONLY: update_tile_halo_left_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max,n_xmin,n_xmax,n_ymin,n_ymax,n_zmin,n_zmax,depth
    INTEGER(C_INT) :: fields(19)
    REAL(C_DOUBLE) :: density0(*),energy0(*),pressure(*),viscosity(*),soundspeed(*),density1(*),energy1(*)
    REAL(C_DOUBLE) :: xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*),vol_flux_x(*)
    REAL(C_DOUBLE) :: vol_flux_y(*),vol_flux_z(*),mass_flux_x(*),mass_flux_y(*),mass_flux_z(*)
    REAL(C_DOUBLE) :: n_density0(*),n_energy0(*),n_pressure(*),n_viscosity(*),n_soundspeed(*),n_density1(*),n_energy1(*)
    REAL(C_DOUBLE) :: n_xvel0(*),n_yvel0(*),n_zvel0(*),n_xvel1(*),n_yvel1(*),n_zvel1(*),n_vol_flux_x(*)
    REAL(C_DOUBLE) :: n_vol_flux_y(*),n_vol_flux_z(*),n_mass_flux_x(*),n_mass_flux_y(*),n_mass_flux_z(*)
    INTEGER :: idepth
    idepth = depth
    CALL update_tile_halo_left_kernel( &
         x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0,pressure,viscosity,soundspeed, &
         density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y, &
         vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,n_xmin,n_xmax,n_ymin,n_ymax, &
         n_zmin,n_zmax,n_density0,n_energy0,n_pressure,n_viscosity,n_soundspeed, &
         n_density1,n_energy1,n_xvel0,n_yvel0,n_zvel0,n_xvel1,n_yvel1,n_zvel1, &
         n_vol_flux_x,n_vol_flux_y,n_vol_flux_z,n_mass_flux_x,n_mass_flux_y,n_mass_flux_z, &
         fields,idepth)
  END SUBROUTINE ref_update_tile_halo_left_kernel

  SUBROUTINE ref_update_tile_halo_right_kernel( &
      x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0,pressure,viscosity,soundspeed, &
      density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y, &
      vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,n_xmin,n_xmax,n_ymin,n_ymax,n_zmin, &
      n_zmax,n_density0,n_energy0,n_pressure,n_viscosity,n_soundspeed,n_density1, &
      n_energy1,n_xvel0,n_yvel0,n_zvel0,n_xvel1,n_yvel1,n_zvel1,n_vol_flux_x,n_vol_flux_y, &
      n_vol_flux_z,n_mass_flux_x,n_mass_flux_y,n_mass_flux_z,fields,depth) BIND(C,NAME='ref_update_tile_halo_right_kernel')
    USE update_tile_halo_kernel_module, ONLY: update_tile_halo_right_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max,n_xmin,n_xmax,n_ymin,n_ymax,n_zmin,n_zmax,depth
    INTEGER(C_INT) :: fields(19)
    REAL(C_DOUBLE) :: density0(*),energy0(*),pressure(*),viscosity(*),soundspeed(*),density1(*),energy1(*)
    REAL(C_DOUBLE) :: xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*),vol_flux_x(*)
    REAL(C_DOUBLE) :: vol_flux_y(*),vol_flux_z(*),mass_flux_x(*),mass_flux_y(*),mass_flux_z(*)
    REAL(C_DOUBLE) :: n_density0(*),n_energy0(*),n_pressure(*),n_viscosity(*),n_soundspeed(*),n_density1(*),n_energy1(*)
    REAL(C_DOUBLE) :: n_xvel0(*),n_yvel0(*),n_zvel0(*),n_xvel1(*),n_yvel1(*),n_zvel1(*),n_vol_flux_x(*)
    REAL(C_DOUBLE) :: n_vol_flux_y(*),n_vol_flux_z(*),n_mass_flux_x(*),n_mass_flux_y(*),n_mass_flux_z(*)
    INTEGER :: idepth
    idepth = depth
    CALL update_tile_halo_right_kernel( &
         x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0,pressure,viscosity,soundspeed, &
         density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y, &
         vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,n_xmin,n_xmax,n_ymin,n_ymax, &
         n_zmin,n_zmax,n_density0,n_energy0,n_pressure,n_viscosity,n_soundspeed, &
         n_density1,n_energy1,n_xvel0,n_yvel0,n_zvel0,n_xvel1,n_yvel1,n_zvel1, &
         n_vol_flux_x,n_vol_flux_y,n_vol_flux_z,n_mass_flux_x,n_mass_flux_y,n_mass_flux_z, &
         fields,idepth)
  END SUBROUTINE ref_update_tile_halo_right_kernel

  SUBROUTINE ref_update_tile_halo_top_kernel( &
      x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0,pressure,viscosity,soundspeed, &
      density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y, &
      vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,n_xmin,n_xmax,n_ymin,n_ymax,n_zmin, &
      n_zmax,n_density0,n_energy0,n_pressure,n_viscosity,n_soundspeed,n_density1, &
      n_energy1,n_xvel0,n_yvel0,n_zvel0,n_xvel1,n_yvel1,n_zvel1,n_vol_flux_x,n_vol_flux_y, &
      n_vol_flux_z,n_mass_flux_x,n_mass_flux_y,n_mass_flux_z,fields,depth) BIND(C,NAME='ref_update_tile_halo_top_kernel')
    USE update_tile_halo_kernel_module, ONLY: update_tile_halo_top_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max,n_xmin,n_xmax,n_ymin,n_ymax,n_zmin,n_zmax,depth
    INTEGER(C_INT) :: fields(19)
    REAL(C_DOUBLE) :: density0(*),energy0(*),pressure(*),viscosity(*),soundspeed(*),density1(*),energy1(*)
    REAL(C_DOUBLE) :: xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*),vol_flux_x(*)
    REAL(C_DOUBLE) :: vol_flux_y(*),vol_flux_z(*),mass_flux_x(*),mass_flux_y(*),mass_flux_z(*)
    REAL(C_DOUBLE) :: n_density0(*),n_energy0(*),n_pressure(*),n_viscosity(*),n_soundspeed(*),n_density1(*),n_energy1(*)
    REAL(C_DOUBLE) :: n_xvel0(*),n_yvel0(*),n_zvel0(*),n_xvel1(*),n_yvel1(*),n_zvel1(*),n_vol_flux_x(*)
    REAL(C_DOUBLE) :: n_vol_flux_y(*),n_vol_flux_z(*),n_mass_flux_x(*),n_mass_flux_y(*),n_mass_flux_z(*)
    INTEGER :: idepth
    idepth = depth
    CALL update_tile_halo_top_kernel( &
         x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0,pressure,viscosity,soundspeed, &
         density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y, &
         vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,n_xmin,n_xmax,n_ymin,n_ymax, &
         n_zmin,n_zmax,n_density0,n_energy0,n_pressure,n_viscosity,n_soundspeed, &
         n_density1,n_energy1,n_xvel0,n_yvel0,n_zvel0,n_xvel1,n_yvel1,n_zvel1, &
         n_vol_flux_x,n_vol_flux_y,n_vol_flux_z,n_mass_flux_x,n_mass_flux_y,n_mass_flux_z, &
         fields,idepth)
  END SUBROUTINE ref_update_tile_halo_top_kernel

  SUBROUTINE ref_update_tile_halo_bottom_kernel( &
      x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0,pressure,viscosity,soundspeed, &
      density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y, &
      vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,n_xmin,n_xmax,n_ymin,n_ymax,n_zmin, &
      n_zmax,n_density0,n_energy0,n_pressure,n_viscosity,n_soundspeed,n_density1, &
      n_energy1,n_xvel0,n_yvel0,n_zvel0,n_xvel1,n_yvel1,n_zvel1,n_vol_flux_x,n_vol_flux_y, &
      n_vol_flux_z,n_mass_flux_x,n_mass_flux_y,n_mass_flux_z,fields,depth) BIND(C,NAME='ref_update_tile_halo_bottom_kernel')
    USE update_tile_halo_kernel_module, ONLY: update_tile_halo_bottom_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max,n_xmin,n_xmax,n_ymin,n_ymax,n_zmin,n_zmax,depth
    INTEGER(C_INT) :: fields(19)
    REAL(C_DOUBLE) :: density0(*),energy0(*),pressure(*),viscosity(*),soundspeed(*),density1(*),energy1(*)
    REAL(C_DOUBLE) :: xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*),vol_flux_x(*)
    REAL(C_DOUBLE) :: vol_flux_y(*),vol_flux_z(*),mass_flux_x(*),mass_flux_y(*),mass_flux_z(*)
    REAL(C_DOUBLE) :: n_density0(*),n_energy0(*),n_pressure(*),n_viscosity(*),n_soundspeed(*),n_density1(*),n_energy1(*)
    REAL(C_DOUBLE) :: n_xvel0(*),n_yvel0(*),n_zvel0(*),n_xvel1(*),n_yvel1(*),n_zvel1(*),n_vol_flux_x(*)
    REAL(C_DOUBLE) :: n_vol_flux_y(*),n_vol_flux_z(*),n_mass_flux_x(*),n_mass_flux_y(*),n_mass_flux_z(*)
    INTEGER :: idepth
    idepth = depth
    CALL update_tile_halo_bottom_kernel( &
         x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0,pressure,viscosity,soundspeed, &
         density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y, &
         vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,n_xmin,n_xmax,n_ymin,n_ymax, &
         n_zmin,n_zmax,n_density0,n_energy0,n_pressure,n_viscosity,n_soundspeed, &
         n_density1,n_energy1,n_xvel0,n_yvel0,n_zvel0,n_xvel1,n_yvel1,n_zvel1, &
         n_vol_flux_x,n_vol_flux_y,n_vol_flux_z,n_mass_flux_x,n_mass_flux_y,n_mass_flux_z, &
         fields,idepth)
  END SUBROUTINE ref_update_tile_halo_bottom_kernel

  SUBROUTINE ref_update_tile_halo_front_kernel( &
      x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0,pressure,viscosity,soundspeed, &
      density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y, &
      vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,n_xmin,n_xmax,n_ymin,n_ymax,n_zmin, &
      n_zmax,n_density0,n_energy0,n_pressure,n_viscosity,n_soundspeed,n_density1, &
      n_energy1,n_xvel0,n_yvel0,n_zvel0,n_xvel1,n_yvel1,n_zvel1,n_vol_flux_x,n_vol_flux_y, &
      n_vol_flux_z,n_mass_flux_x,n_mass_flux_y,n_mass_flux_z,fields,depth) BIND(C,NAME='ref_update_tile_halo_front_kernel')
    USE update_tile_halo_kernel_module, ONLY: update_tile_halo_front_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max,n_xmin,n_xmax,n_ymin,n_ymax,n_zmin,n_zmax,depth
    INTEGER(C_INT) :: fields(19)
    REAL(C_DOUBLE) :: density0(*),energy0(*),pressure(*),viscosity(*),soundspeed(*),density1(*),energy1(*)
    REAL(C_DOUBLE) :: xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*),vol_flux_x(*)
    REAL(C_DOUBLE) :: vol_flux_y(*),vol_flux_z(*),mass_flux_x(*),mass_flux_y(*),mass_flux_z(*)
    REAL(C_DOUBLE) :: n_density0(*),n_energy0(*),n_pressure(*),n_viscosity(*),n_soundspeed(*),n_density1(*),n_energy1(*)
    REAL(C_DOUBLE) :: n_xvel0(*),n_yvel0(*),n_zvel0(*),n_xvel1(*),n_yvel1(*),n_zvel1(*),n_vol_flux_x(*)
    REAL(C_DOUBLE) :: n_vol_flux_y(*),n_vol_flux_z(*),n_mass_flux_x(*),n_mass_flux_y(*),n_mass_flux_z(*)
    INTEGER :: idepth
    idepth = depth
    CALL update_tile_halo_front_kernel( &
         x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0,pressure,viscosity,soundspeed, &
         density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y, &
         vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,n_xmin,n_xmax,n_ymin,n_ymax, &
         n_zmin,n_zmax,n_density0,n_energy0,n_pressure,n_viscosity,n_soundspeed, &
         n_density1,n_energy1,n_xvel0,n_yvel0,n_zvel0,n_xvel1,n_yvel1,n_zvel1, &
         n_vol_flux_x,n_vol_flux_y,n_vol_flux_z,n_mass_flux_x,n_mass_flux_y,n_mass_flux_z, &
         fields,idepth)
  END SUBROUTINE ref_update_tile_halo_front_kernel

  SUBROUTINE ref_update_tile_halo_back_kernel( &
      x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0,pressure,viscosity,soundspeed, &
      density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y, &
      vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,n_xmin,n_xmax,n_ymin,n_ymax,n_zmin, &
      n_zmax,n_density0,n_energy0,n_pressure,n_viscosity,n_soundspeed,n_density1, &
      n_energy1,n_xvel0,n_yvel0,n_zvel0,n_xvel1,n_yvel1,n_zvel1,n_vol_flux_x,n_vol_flux_y, &
      n_vol_flux_z,n_mass_flux_x,n_mass_flux_y,n_mass_flux_z,fields,depth) BIND(C,NAME='ref_update_tile_halo_back_kernel')
    USE update_tile_halo_kernel_module, ONLY: update_tile_halo_back_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max,n_xmin,n_xmax,n_ymin,n_ymax,n_zmin,n_zmax,depth
    INTEGER(C_INT) :: fields(19)
    REAL(C_DOUBLE) :: density0(*),energy0(*),pressure(*),viscosity(*),soundspeed(*),density1(*),energy1(*)
    REAL(C_DOUBLE) :: xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*),vol_flux_x(*)
    REAL(C_DOUBLE) :: vol_flux_y(*),vol_flux_z(*),mass_flux_x(*),mass_flux_y(*),mass_flux_z(*)
    REAL(C_DOUBLE) :: n_density0(*),n_energy0(*),n_pressure(*),n_viscosity(*),n_soundspeed(*),n_density1(*),n_energy1(*)
    REAL(C_DOUBLE) :: n_xvel0(*),n_yvel0(*),n_zvel0(*),n_xvel1(*),n_yvel1(*),n_zvel1(*),n_vol_flux_x(*)
    REAL(C_DOUBLE) :: n_vol_flux_y(*),n_vol_flux_z(*),n_mass_flux_x(*),n_mass_flux_y(*),n_mass_flux_z(*)
    INTEGER :: idepth
    idepth = depth
    CALL update_tile_halo_back_kernel( &
         x_min,x_max,y_min,y_max,z_min,z_max,density0,energy0,pressure,viscosity,soundspeed, &
         density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y, &
         vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,n_xmin,n_xmax,n_ymin,n_ymax, &
         n_zmin,n_zmax,n_density0,n_energy0,n_pressure,n_viscosity,n_soundspeed, &
         n_density1,n_energy1,n_xvel0,n_yvel0,n_zvel0,n_xvel1,n_yvel1,n_zvel1, &
         n_vol_flux_x,n_vol_flux_y,n_vol_flux_z,n_mass_flux_x,n_mass_flux_y,n_mass_flux_z, &
         fields,idepth)
  END SUBROUTINE ref_update_tile_halo_back_kernel

END MODULE ref_bind_tile
