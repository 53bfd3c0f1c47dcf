! C-callable wrappers around the reference's own *_kernel subroutines.
!
! TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together
! with the reference's kernel sources, read in place from the reference
! tree, into oracle/_ref/libclover_ref.so.  The oracle driver and the tests
! call the reference kernels through these wrappers to pin the C
! restatement.  Scalars arrive by value, arrays as base addresses of
! Fortran column-major storage (assumed size here, sequence-associated with
! the kernels' explicit-shape dummies).  LOGICAL arguments arrive as C int.

MODULE ref_bind
  USE ISO_C_BINDING
  IMPLICIT NONE
CONTAINS

  SUBROUTINE ref_ideal_gas_kernel(x_min,x_max,y_min,y_max,z_min,z_max,density,energy,pressure,soundspeed) &
      BIND(C,NAME='ref_ideal_gas_kernel')
    USE ideal_gas_kernel_module, ONLY: ideal_gas_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(C_DOUBLE) :: density(*),energy(*),pressure(*),soundspeed(*)
    CALL ideal_gas_kernel(x_min,x_max,y_min,y_max,z_min,z_max,density,energy,pressure,soundspeed)
  END SUBROUTINE ref_ideal_gas_kernel

  SUBROUTINE ref_viscosity_kernel(x_min,x_max,y_min,y_max,z_min,z_max,xarea,yarea,zarea, &
      celldx,celldy,celldz,density0,pressure,viscosity,xvel0,yvel0,zvel0)              &
      BIND(C,NAME='ref_viscosity_kernel')
    USE viscosity_kernel_module, ONLY: viscosity_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(C_DOUBLE) :: xarea(*),yarea(*),zarea(*),celldx(*),celldy(*),celldz(*)
    REAL(C_DOUBLE) :: density0(*),pressure(*),viscosity(*),xvel0(*),yvel0(*),zvel0(*)
    CALL viscosity_kernel(x_min,x_max,y_min,y_max,z_min,z_max,xarea,yarea,zarea, &
                          celldx,celldy,celldz,density0,pressure,viscosity,xvel0,yvel0,zvel0)
  END SUBROUTINE ref_viscosity_kernel

  SUBROUTINE ref_calc_dt_kernel(x_min,x_max,y_min,y_max,z_min,z_max,g_small,g_big,dtmin,      &
      dtc_safe,dtu_safe,dtv_safe,dtw_safe,dtdiv_safe,xarea,yarea,zarea,cellx,celly,cellz,     &
      celldx,celldy,celldz,volume,density0,energy0,pressure,viscosity_a,soundspeed,           &
      xvel0,yvel0,zvel0,dt_min,dt_min_val,dtl_control,xl_pos,yl_pos,zl_pos,jldt,kldt,lldt,small) &
      BIND(C,NAME='ref_calc_dt_kernel')
    USE calc_dt_kernel_module, ONLY: calc_dt_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(C_DOUBLE), VALUE :: g_small,g_big,dtmin,dtc_safe,dtu_safe,dtv_safe,dtw_safe,dtdiv_safe
    REAL(C_DOUBLE) :: xarea(*),yarea(*),zarea(*),cellx(*),celly(*),cellz(*)
    REAL(C_DOUBLE) :: celldx(*),celldy(*),celldz(*),volume(*),density0(*),energy0(*)
    REAL(C_DOUBLE) :: pressure(*),viscosity_a(*),soundspeed(*),xvel0(*),yvel0(*),zvel0(*),dt_min(*)
    REAL(C_DOUBLE) :: dt_min_val,xl_pos,yl_pos,zl_pos
    INTEGER(C_INT) :: dtl_control,jldt,kldt,lldt,small
    CALL calc_dt_kernel(x_min,x_max,y_min,y_max,z_min,z_max,g_small,g_big,dtmin,          &
                        dtc_safe,dtu_safe,dtv_safe,dtw_safe,dtdiv_safe,xarea,yarea,zarea, &
                        cellx,celly,cellz,celldx,celldy,celldz,volume,density0,energy0,   &
                        pressure,viscosity_a,soundspeed,xvel0,yvel0,zvel0,dt_min,         &
                        dt_min_val,dtl_control,xl_pos,yl_pos,zl_pos,jldt,kldt,lldt,small)
  END SUBROUTINE ref_calc_dt_kernel

  SUBROUTINE ref_pdv_kernel(predict,x_min,x_max,y_min,y_max,z_min,z_max,dt,xarea,yarea,zarea,volume, &
      density0,density1,energy0,energy1,pressure,viscosity,xvel0,xvel1,yvel0,yvel1,zvel0,zvel1)      &
      BIND(C,NAME='ref_pdv_kernel')
    USE PdV_kernel_module, ONLY: PdV_kernel
    INTEGER(C_INT), VALUE :: predict,x_min,x_max,y_min,y_max,z_min,z_max
    REAL(C_DOUBLE), VALUE :: dt
    REAL(C_DOUBLE) :: xarea(*),yarea(*),zarea(*),volume(*),density0(*),density1(*),energy0(*),energy1(*)
    REAL(C_DOUBLE) :: pressure(*),viscosity(*),xvel0(*),xvel1(*),yvel0(*),yvel1(*),zvel0(*),zvel1(*)
    LOGICAL :: lpredict
    lpredict = (predict /= 0)
    CALL PdV_kernel(lpredict,x_min,x_max,y_min,y_max,z_min,z_max,dt,xarea,yarea,zarea,volume, &
                    density0,density1,energy0,energy1,pressure,viscosity,                    &
                    xvel0,xvel1,yvel0,yvel1,zvel0,zvel1)
  END SUBROUTINE ref_pdv_kernel

  SUBROUTINE ref_revert_kernel(x_min,x_max,y_min,y_max,z_min,z_max,density0,density1,energy0,energy1) &
      BIND(C,NAME='ref_revert_kernel')
    USE revert_kernel_module, ONLY: revert_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(C_DOUBLE) :: density0(*),density1(*),energy0(*),energy1(*)
    CALL revert_kernel(x_min,x_max,y_min,y_max,z_min,z_max,density0,density1,energy0,energy1)
  END SUBROUTINE ref_revert_kernel

  SUBROUTINE ref_accelerate_kernel(x_min,x_max,y_min,y_max,z_min,z_max,dt,xarea,yarea,zarea,volume, &
      density0,pressure,viscosity,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1)                             &
      BIND(C,NAME='ref_accelerate_kernel')
    USE accelerate_kernel_module, ONLY: accelerate_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(C_DOUBLE), VALUE :: dt
    REAL(C_DOUBLE) :: xarea(*),yarea(*),zarea(*),volume(*),density0(*),pressure(*),viscosity(*)
    REAL(C_DOUBLE) :: xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*)
    CALL accelerate_kernel(x_min,x_max,y_min,y_max,z_min,z_max,dt,xarea,yarea,zarea,volume, &
                           density0,pressure,viscosity,xvel0,yvel0,zvel0,xvel1,yvel1,zvel1)
  END SUBROUTINE ref_accelerate_kernel

  SUBROUTINE ref_flux_calc_kernel(x_min,x_max,y_min,y_max,z_min,z_max,dt,xarea,yarea,zarea, &
      xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y,vol_flux_z)                &
      BIND(C,NAME='ref_flux_calc_kernel')
    USE flux_calc_kernel_module, ONLY: flux_calc_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(C_DOUBLE), VALUE :: dt
    REAL(C_DOUBLE) :: xarea(*),yarea(*),zarea(*),xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*)
    REAL(C_DOUBLE) :: vol_flux_x(*),vol_flux_y(*),vol_flux_z(*)
    CALL flux_calc_kernel(x_min,x_max,y_min,y_max,z_min,z_max,dt,xarea,yarea,zarea, &
                          xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y,vol_flux_z)
  END SUBROUTINE ref_flux_calc_kernel

  SUBROUTINE ref_advec_cell_kernel(x_min,x_max,y_min,y_max,z_min,z_max,advect_x,dir,sweep_number, &
      vertexdx,vertexdy,vertexdz,volume,density1,energy1,mass_flux_x,vol_flux_x,mass_flux_y,       &
      vol_flux_y,mass_flux_z,vol_flux_z,pre_vol,post_vol,pre_mass,post_mass,advec_vol,post_ener,   &
      ener_flux) BIND(C,NAME='ref_advec_cell_kernel')
    USE advec_cell_kernel_module, ONLY: advec_cell_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max,advect_x,dir,sweep_number
    REAL(C_DOUBLE) :: vertexdx(*),vertexdy(*),vertexdz(*),volume(*),density1(*),energy1(*)
    REAL(C_DOUBLE) :: mass_flux_x(*),vol_flux_x(*),mass_flux_y(*),vol_flux_y(*),mass_flux_z(*),vol_flux_z(*)
    REAL(C_DOUBLE) :: pre_vol(*),post_vol(*),pre_mass(*),post_mass(*),advec_vol(*),post_ener(*),ener_flux(*)
    LOGICAL :: ladvect_x
    INTEGER :: idir,isweep
    ladvect_x = (advect_x /= 0)
    idir = dir
    isweep = sweep_number
    CALL advec_cell_kernel(x_min,x_max,y_min,y_max,z_min,z_max,ladvect_x,idir,isweep,          &
                           vertexdx,vertexdy,vertexdz,volume,density1,energy1,                &
                           mass_flux_x,vol_flux_x,mass_flux_y,vol_flux_y,mass_flux_z,vol_flux_z, &
                           pre_vol,post_vol,pre_mass,post_mass,advec_vol,post_ener,ener_flux)
  END SUBROUTINE ref_advec_cell_kernel

  SUBROUTINE ref_advec_mom_kernel(x_min,x_max,y_min,y_max,z_min,z_max,vel1,mass_flux_x,vol_flux_x, &
      mass_flux_y,vol_flux_y,mass_flux_z,vol_flux_z,volume,density1,node_flux,node_mass_post,       &
      node_mass_pre,advec_vel,mom_flux,pre_vol,post_vol,celldx,celldy,celldz,advect_x,which_vel,    &
      sweep_number,direction) BIND(C,NAME='ref_advec_mom_kernel')
    USE advec_mom_kernel_mod, ONLY: advec_mom_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max,advect_x,which_vel,sweep_number,direction
    REAL(C_DOUBLE) :: vel1(*),mass_flux_x(*),vol_flux_x(*),mass_flux_y(*),vol_flux_y(*)
    REAL(C_DOUBLE) :: mass_flux_z(*),vol_flux_z(*),volume(*),density1(*),node_flux(*),node_mass_post(*)
    REAL(C_DOUBLE) :: node_mass_pre(*),advec_vel(*),mom_flux(*),pre_vol(*),post_vol(*)
    REAL(C_DOUBLE) :: celldx(*),celldy(*),celldz(*)
    LOGICAL :: ladvect_x
    INTEGER :: iwhich,isweep,idir
    ladvect_x = (advect_x /= 0)
    iwhich = which_vel
    isweep = sweep_number
    idir = direction
    CALL advec_mom_kernel(x_min,x_max,y_min,y_max,z_min,z_max,vel1,mass_flux_x,vol_flux_x,     &
                          mass_flux_y,vol_flux_y,mass_flux_z,vol_flux_z,volume,density1,       &
                          node_flux,node_mass_post,node_mass_pre,advec_vel,mom_flux,pre_vol,   &
                          post_vol,celldx,celldy,celldz,ladvect_x,iwhich,isweep,idir)
  END SUBROUTINE ref_advec_mom_kernel

  SUBROUTINE ref_reset_field_kernel(x_min,x_max,y_min,y_max,z_min,z_max,density0,density1,energy0, &
      energy1,xvel0,xvel1,yvel0,yvel1,zvel0,zvel1) BIND(C,NAME='ref_reset_field_kernel')
    USE reset_field_kernel_module, ONLY: reset_field_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(C_DOUBLE) :: density0(*),density1(*),energy0(*),energy1(*)
    REAL(C_DOUBLE) :: xvel0(*),xvel1(*),yvel0(*),yvel1(*),zvel0(*),zvel1(*)
    CALL reset_field_kernel(x_min,x_max,y_min,y_max,z_min,z_max,density0,density1,energy0,energy1, &
                            xvel0,xvel1,yvel0,yvel1,zvel0,zvel1)
  END SUBROUTINE ref_reset_field_kernel

  SUBROUTINE ref_update_halo_kernel(x_min,x_max,y_min,y_max,z_min,z_max,chunk_neighbours,tile_neighbours, &
      density0,energy0,pressure,viscosity,soundspeed,density1,energy1,xvel0,yvel0,zvel0,xvel1,yvel1,     &
      zvel1,vol_flux_x,vol_flux_y,vol_flux_z,mass_flux_x,mass_flux_y,mass_flux_z,fields,depth)          &
      BIND(C,NAME='ref_update_halo_kernel')
    USE update_halo_kernel_module, ONLY: update_halo_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max,depth
    INTEGER(C_INT) :: chunk_neighbours(6),tile_neighbours(6),fields(19)
    REAL(C_DOUBLE) :: density0(*),energy0(*),pressure(*),viscosity(*),soundspeed(*),density1(*),energy1(*)
    REAL(C_DOUBLE) :: xvel0(*),yvel0(*),zvel0(*),xvel1(*),yvel1(*),zvel1(*)
    REAL(C_DOUBLE) :: vol_flux_x(*),vol_flux_y(*),vol_flux_z(*),mass_flux_x(*),mass_flux_y(*),mass_flux_z(*)
    INTEGER :: idepth
    idepth = depth
    CALL update_halo_kernel(x_min,x_max,y_min,y_max,z_min,z_max,chunk_neighbours,tile_neighbours, &
                            density0,energy0,pressure,viscosity,soundspeed,density1,energy1,      &
                            xvel0,yvel0,zvel0,xvel1,yvel1,zvel1,vol_flux_x,vol_flux_y,vol_flux_z, &
                            mass_flux_x,mass_flux_y,mass_flux_z,fields,idepth)
  END SUBROUTINE ref_update_halo_kernel

  SUBROUTINE ref_field_summary_kernel(x_min,x_max,y_min,y_max,z_min,z_max,volume,density0,energy0, &
      pressure,xvel0,yvel0,zvel0,vol,mass,ie,ke,press) BIND(C,NAME='ref_field_summary_kernel')
    USE field_summary_kernel_module, ONLY: field_summary_kernel
    INTEGER(C_INT), VALUE :: x_min,x_max,y_min,y_max,z_min,z_max
    REAL(C_DOUBLE) :: volume(*),density0(*),energy0(*),pressure(*),xvel0(*),yvel0(*),zvel0(*)
    REAL(C_DOUBLE) :: vol,mass,ie,ke,press
    CALL field_summary_kernel(x_min,x_max,y_min,y_max,z_min,z_max,volume,density0,energy0,pressure, &
                              xvel0,yvel0,zvel0,vol,mass,ie,ke,press)
  END SUBROUTINE ref_field_summary_kernel

END MODULE ref_bind
