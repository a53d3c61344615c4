! Driver program for the reference Fortran model (oracle/_ref/msw_ref).
!
! Test infrastructure only: it links the reference sources compiled where they
! lie (src/mpas_sw_module.f90 and the dyncore/tlm/adj/operators files) and
! drives the reference module exactly as src/mpas_sw_driver.py does:
! set dimensions (read_dims, :33-61), allocate and fill the mesh (read_vars,
! :64-119), run the test case (initial_conditions, :220-234), allocate the
! TLM/ADJ arrays, sw_mpas_init_block, then step.  Inputs and outputs are flat
! binary files written and read by oracle/ref.py.
!
! usage: msw_ref <input.bin> <mode> <nsteps> <output.bin>
!   modes: ic fwd tlm adj diag diagtl tendadj time_fwd time_tlm time_adj
program msw_ref
  use mpas_sw_module
  implicit none

  character(len=512) :: fin, fout, mode, arg
  integer :: nsteps, hdr(16), ios, k, istep
  real(kind=RKIND) :: cfg(8)
  integer :: has_state, has_tl, has_ad
  real(kind=RKIND), allocatable :: st_h(:), st_u(:), tl_h(:), tl_u(:), ad_h(:), ad_u(:)
  real(kind=RKIND), allocatable :: traj_h(:,:), traj_u(:,:), usrc_in(:), rseed(:,:)
  real(kind=RKIND), allocatable :: ds_e(:,:), ds_v(:,:), ds_c(:,:)
  integer(8) :: t0, t1, rate
  integer, parameter :: uin = 21, uout = 22

  call get_command_argument(1, fin)
  call get_command_argument(2, mode)
  call get_command_argument(3, arg)
  read(arg, *) nsteps
  call get_command_argument(4, fout)

  open(uin, file=trim(fin), access='stream', form='unformatted', status='old', iostat=ios)
  if (ios /= 0) stop 'cannot open input'
  read(uin) hdr
  read(uin) cfg
  if (hdr(1) /= 1297307462) stop 'bad magic'

  ! ---- dimensions, as read_dims does
  nVertLevels = 1
  nCells = hdr(3); nEdges = hdr(4); nVertices = hdr(5)
  maxEdges = hdr(6); maxEdges2 = hdr(7); vertexDegree = hdr(8)
  R3 = 3
  nTracers = 1
  nCellsSolve = nCells; nEdgesSolve = nEdges
  cellStart = 1; cellEnd = nCells; vertexStart = 1; vertexEnd = nVertices
  edgeStart = 1; edgeEnd = nEdges
  cellSolveStart = 1; cellSolveEnd = nCells; edgeSolveStart = 1; edgeSolveEnd = nEdges

  ! ---- configuration, as read_configs does
  config_test_case = hdr(9)
  config_thickness_adv_order = hdr(10)
  config_monotonic = hdr(11) /= 0
  config_wind_stress = hdr(12) /= 0
  config_bottom_drag = hdr(13) /= 0
  has_state = hdr(14); has_tl = hdr(15); has_ad = hdr(16)
  ! cfg(8) /= 0: config_h_ScaleWithMesh, which src/mpas_sw_driver.py cannot
  ! reach (mixed case through f2py) but compute_mesh_scaling reads
  config_h_ScaleWithMesh = cfg(8) /= 0.0
  config_dt = cfg(1)
  config_apvm_upwinding = cfg(2)
  config_h_mom_eddy_visc2 = cfg(3)
  config_h_mom_eddy_visc4 = cfg(4)

  call var_allocation()
  on_a_sphere = .true.
  is_periodic = .false.
  sphere_radius = 1.0

  read(uin) latCell, lonCell, xCell, yCell, zCell
  read(uin) latEdge, lonEdge, xEdge, yEdge, zEdge
  read(uin) latVertex, lonVertex, xVertex, yVertex, zVertex
  read(uin) cellsOnEdge, nEdgesOnCell, nEdgesOnEdge, edgesOnCell, edgesOnEdge
  read(uin) weightsOnEdge, dvEdge, dcEdge, angleEdge, areaCell, areaTriangle
  read(uin) cellsOnCell, verticesOnCell, verticesOnEdge, edgesOnVertex, cellsOnVertex
  read(uin) kiteAreasOnVertex, meshDensity
  allocate(st_h(nCells), st_u(nEdges), tl_h(nCells), tl_u(nEdges), ad_h(nCells), ad_u(nEdges))
  st_h = 0; st_u = 0; tl_h = 0; tl_u = 0; ad_h = 0; ad_u = 0
  if (has_state /= 0) read(uin) st_h, st_u
  if (has_tl /= 0) read(uin) tl_h, tl_u
  if (has_ad /= 0) read(uin) ad_h, ad_u
  ! cfg(5) /= 0: a wind-stress source field u_src follows (module array,
  ! src/mpas_sw_module.f90:78,223-224; zero unless the caller sets it)
  if (cfg(5) /= 0.0) then
     allocate(usrc_in(nEdges))
     read(uin) usrc_in
  end if
  ! cfg(6) /= 0: seeds of the reconstructed-velocity adjoints follow
  ! (uReconstructX/Y/Z/Zonal/Meridional_ad, src/mpas_sw_module.f90:168-172),
  ! consumed by the first sw_rk4_adj
  if (cfg(6) /= 0.0) then
     allocate(rseed(nCells, 5))
     read(uin) rseed
  end if
  ! cfg(7) /= 0: seeds of the diagnostic adjoints follow, in the order v,
  ! h_edge, h_vertex, circulation, vorticity, divergence, ke, pv_edge,
  ! pv_vertex, pv_cell, vorticity_cell, gradPVn, gradPVt (module arrays,
  ! src/mpas_sw_module.f90:450-463), consumed by the first sw_rk4_adj
  if (cfg(7) /= 0.0) then
     allocate(ds_e(nEdges, 5), ds_v(nVertices, 4), ds_c(nCells, 4))
     read(uin) ds_e(:,1), ds_e(:,2), ds_v(:,1), ds_v(:,2), ds_v(:,3), ds_c(:,1), ds_c(:,2), &
          ds_e(:,3), ds_v(:,4), ds_c(:,3), ds_c(:,4), ds_e(:,4), ds_e(:,5)
  end if
  close(uin)

  ! ---- initial conditions, as initial_conditions does
  select case (config_test_case)
  case (1)
     call sw_test_case_1()
  case (2)
     call sw_test_case_2()
  case (5)
     call sw_test_case_5()
  case (6)
     call sw_test_case_6()
  case default
     stop 'unsupported test case'
  end select
  if (has_state /= 0) then
     h(1,1,:) = st_h
     u(1,1,:) = st_u
  end if
  if (allocated(usrc_in)) u_src(1,:) = usrc_in
  call var_allocation_tlm()
  call var_allocation_adj()
  call sw_mpas_init_block()
  if (allocated(rseed)) then
     uReconstructX_ad(1,:) = rseed(:,1)
     uReconstructY_ad(1,:) = rseed(:,2)
     uReconstructZ_ad(1,:) = rseed(:,3)
     uReconstructZonal_ad(1,:) = rseed(:,4)
     uReconstructMeridional_ad(1,:) = rseed(:,5)
  end if
  if (allocated(ds_e)) then
     v_ad(1,:) = ds_e(:,1)
     h_edge_ad(1,:) = ds_e(:,2)
     h_vertex_ad(1,:) = ds_v(:,1)
     circulation_ad(1,:) = ds_v(:,2)
     vorticity_ad(1,:) = ds_v(:,3)
     divergence_ad(1,:) = ds_c(:,1)
     ke_ad(1,:) = ds_c(:,2)
     pv_edge_ad(1,:) = ds_e(:,3)
     pv_vertex_ad(1,:) = ds_v(:,4)
     pv_cell_ad(1,:) = ds_c(:,3)
     vorticity_cell_ad(1,:) = ds_c(:,4)
     gradPVn_ad(1,:) = ds_e(:,4)
     gradPVt_ad(1,:) = ds_e(:,5)
  end if

  open(uout, file=trim(fout), access='stream', form='unformatted', status='replace')

  select case (trim(mode))
  case ('ic')
     call dump_ic()
     call dump_diag()
  case ('fwd')
     do istep = 1, nsteps
        call sw_rk4()
     end do
     call dump_state()
     call dump_diag()
     call dump_recon()
  case ('tlm')
     h_tl(1,1,:) = tl_h
     u_tl(1,1,:) = tl_u
     do istep = 1, nsteps
        call sw_rk4_tlm()
     end do
     call dump_state()
     call put('h_tl', h_tl(1,1,:), nCells)
     call put('u_tl', u_tl(1,1,:), nEdges)
     call dump_diag()
     call dump_diag_tl()
     call put('uReconstructZonal_tl', uReconstructZonal_tl(1,:), nCells)
     call put('uReconstructMeridional_tl', uReconstructMeridional_tl(1,:), nCells)
  case ('adj')
     ! forward sweep storing the state at the start of each step, then the
     ! backward sweep feeding that state to sw_rk4_adj (demo/run-adj.py:41-55)
     allocate(traj_h(nCells, max(nsteps,1)), traj_u(nEdges, max(nsteps,1)))
     do istep = 1, nsteps
        traj_h(:,istep) = h(1,1,:)
        traj_u(:,istep) = u(1,1,:)
        call sw_rk4()
     end do
     call put('h_end', h(1,1,:), nCells)
     call put('u_end', u(1,1,:), nEdges)
     h_ad(1,1,:) = ad_h
     u_ad(1,1,:) = ad_u
     do istep = nsteps, 1, -1
        h(1,1,:) = traj_h(:,istep)
        u(1,1,:) = traj_u(:,istep)
        call sw_rk4_adj()
     end do
     call put('h_ad', h_ad(1,1,:), nCells)
     call put('u_ad', u_ad(1,1,:), nEdges)
     call put('h2', h(2,1,:), nCells)
     call put('u2', u(2,1,:), nEdges)
     call put('uReconstructZonal_ad', uReconstructZonal_ad(1,:), nCells)
  case ('diag')
     call sw_compute_solve_diagnostics(config_dt, nCells, nEdges, nVertices, maxEdges, &
          maxEdges2, nVertLevels, vertexDegree, config_monotonic, &
          config_thickness_adv_order, config_apvm_upwinding, &
          h(1,:,:), u(1,:,:), v, vh, h_edge, h_vertex, circulation, vorticity, &
          divergence, ke, pv_edge, pv_vertex, pv_cell, vorticity_cell, &
          gradPVn, gradPVt, weightsOnEdge, kiteAreasOnVertex, cellsOnEdge, &
          cellsOnVertex, cellsOnCell, verticesOnEdge, nEdgesOnCell, edgesOnCell, &
          nEdgesOnEdge, edgesOnEdge, edgesOnVertex, dcEdge, dvEdge, areaCell, &
          areaTriangle, fVertex, deriv_two, boundaryEdge, boundaryCell)
     call sw_compute_tend(nCells, nEdges, nVertices, nVertLevels, &
          maxEdges, maxEdges2, vertexDegree, config_bottom_drag, &
          config_wind_stress,config_h_mom_eddy_visc2, &
          config_h_mom_eddy_visc4, h(1,:,:), u(1,:,:), v, h_edge, circulation, &
          vorticity, divergence, ke, pv_edge, vh, weightsOnEdge, &
          kiteAreasOnVertex, cellsOnEdge, cellsOnVertex, verticesOnEdge, &
          nEdgesOnCell, edgesOnCell, nEdgesOnEdge, edgesOnEdge, edgesOnVertex, &
          dcEdge, dvEdge, areaCell, areaTriangle, h_s, fVertex, fEdge, u_src, &
          meshScalingDel2, meshScalingDel4, tend_h, tend_u)
     call dump_diag()
     call put('tend_h', tend_h(1,:), nCells)
     call put('tend_u', tend_u(1,:), nEdges)
  case ('diagtl')
     h_tl(1,1,:) = tl_h
     u_tl(1,1,:) = tl_u
     call sw_compute_solve_diagnostics_tlm(config_dt, nCells, nEdges, nVertices, maxEdges, &
          maxEdges2, nVertLevels, vertexDegree, config_monotonic, &
          config_thickness_adv_order, config_apvm_upwinding, &
          h(1,:,:), h_tl(1,:,:), u(1,:,:), u_tl(1,:,:), v, v_tl, vh, vh_tl, &
          h_edge, h_edge_tl, h_vertex, h_vertex_tl, circulation, circulation_tl, &
          vorticity, vorticity_tl, divergence, divergence_tl, ke, ke_tl, &
          pv_edge, pv_edge_tl, pv_vertex, pv_vertex_tl, pv_cell, pv_cell_tl, &
          vorticity_cell, vorticity_cell_tl, gradPVn, gradPVn_tl, gradPVt, gradPVt_tl, &
          weightsOnEdge, kiteAreasOnVertex, cellsOnEdge, &
          cellsOnVertex, cellsOnCell, verticesOnEdge, nEdgesOnCell, edgesOnCell, &
          nEdgesOnEdge, edgesOnEdge, edgesOnVertex, dcEdge, dvEdge, areaCell, &
          areaTriangle, fVertex, deriv_two, boundaryEdge, boundaryCell)
     call sw_compute_tend_tlm(nCells, nEdges, nVertices, nVertLevels, &
          maxEdges, maxEdges2, vertexDegree, config_bottom_drag, &
          config_wind_stress,config_h_mom_eddy_visc2, &
          config_h_mom_eddy_visc4, h(1,:,:), h_tl(1,:,:), u(1,:,:), u_tl(1,:,:), &
          v, v_tl, h_edge, h_edge_tl, circulation, circulation_tl, &
          vorticity, vorticity_tl, divergence, divergence_tl, ke, ke_tl, &
          pv_edge, pv_edge_tl, vh, vh_tl, weightsOnEdge, kiteAreasOnVertex, &
          cellsOnEdge, cellsOnVertex, verticesOnEdge, nEdgesOnCell, &
          edgesOnCell, nEdgesOnEdge, edgesOnEdge, edgesOnVertex, &
          dcEdge, dvEdge, areaCell, areaTriangle, h_s, fVertex, fEdge, u_src, &
          meshScalingDel2, meshScalingDel4, &
          tend_h, tend_h_tl, tend_u, tend_u_tl)
     call dump_diag()
     call dump_diag_tl()
     call put('tend_h_tl', tend_h_tl(1,:), nCells)
     call put('tend_u_tl', tend_u_tl(1,:), nEdges)
  case ('tendadj')
     ! adjoint of the full tendency operator at the given state: seeds are the
     ! tendency adjoints, output the state adjoint
     call sw_compute_solve_diagnostics(config_dt, nCells, nEdges, nVertices, maxEdges, &
          maxEdges2, nVertLevels, vertexDegree, config_monotonic, &
          config_thickness_adv_order, config_apvm_upwinding, &
          h(1,:,:), u(1,:,:), v, vh, h_edge, h_vertex, circulation, vorticity, &
          divergence, ke, pv_edge, pv_vertex, pv_cell, vorticity_cell, &
          gradPVn, gradPVt, weightsOnEdge, kiteAreasOnVertex, cellsOnEdge, &
          cellsOnVertex, cellsOnCell, verticesOnEdge, nEdgesOnCell, edgesOnCell, &
          nEdgesOnEdge, edgesOnEdge, edgesOnVertex, dcEdge, dvEdge, areaCell, &
          areaTriangle, fVertex, deriv_two, boundaryEdge, boundaryCell)
     tend_h_ad(1,:) = ad_h
     tend_u_ad(1,:) = ad_u
     h_ad = 0
     u_ad = 0
     call sw_compute_tend_adj(nCells, nEdges, nVertices, nVertLevels, &
          maxEdges, maxEdges2, vertexDegree, config_bottom_drag, &
          config_wind_stress,config_h_mom_eddy_visc2, &
          config_h_mom_eddy_visc4, h(1,:,:), h_ad(1,:,:), &
          u(1,:,:), u_ad(1,:,:), &
          v, v_ad, h_edge, h_edge_ad, circulation, circulation_ad, &
          vorticity, vorticity_ad, divergence, divergence_ad, ke, ke_ad, &
          pv_edge, pv_edge_ad, vh, vh_ad, weightsOnEdge, kiteAreasOnVertex, &
          cellsOnEdge, cellsOnVertex, verticesOnEdge, nEdgesOnCell, &
          edgesOnCell, nEdgesOnEdge, edgesOnEdge, edgesOnVertex, &
          dcEdge, dvEdge, areaCell, areaTriangle, h_s, fVertex, fEdge, u_src, &
          meshScalingDel2, meshScalingDel4, &
          tend_h, tend_h_ad, tend_u, tend_u_ad)
     call sw_compute_solve_diagnostics_adj(config_dt, nCells, nEdges, nVertices, maxEdges, &
          maxEdges2, nVertLevels, vertexDegree, config_monotonic, &
          config_thickness_adv_order, config_apvm_upwinding, &
          h(1,:,:), h_ad(1,:,:), u(1,:,:), u_ad(1,:,:), v, v_ad, vh, vh_ad, &
          h_edge, h_edge_ad, h_vertex, h_vertex_ad, circulation, circulation_ad, &
          vorticity, vorticity_ad, divergence, divergence_ad, ke, ke_ad, &
          pv_edge, pv_edge_ad, pv_vertex, pv_vertex_ad, pv_cell, pv_cell_ad, &
          vorticity_cell, vorticity_cell_ad, gradPVn, gradPVn_ad, gradPVt, gradPVt_ad, &
          weightsOnEdge, kiteAreasOnVertex, cellsOnEdge, &
          cellsOnVertex, cellsOnCell, verticesOnEdge, nEdgesOnCell, edgesOnCell, &
          nEdgesOnEdge, edgesOnEdge, edgesOnVertex, dcEdge, dvEdge, areaCell, &
          areaTriangle, fVertex, deriv_two, boundaryEdge, boundaryCell)
     call put('h_ad', h_ad(1,1,:), nCells)
     call put('u_ad', u_ad(1,1,:), nEdges)
  case ('time_fwd', 'time_tlm', 'time_adj')
     h_tl(1,1,:) = tl_h
     u_tl(1,1,:) = tl_u
     h_ad(1,1,:) = ad_h
     u_ad(1,1,:) = ad_u
     call system_clock(t0, rate)
     do istep = 1, nsteps
        if (trim(mode) == 'time_fwd') then
           call sw_rk4()
        else if (trim(mode) == 'time_tlm') then
           call sw_rk4_tlm()
        else
           call sw_rk4_adj()
        end if
     end do
     call system_clock(t1)
     call put1('seconds', real(t1 - t0, RKIND) / real(rate, RKIND))
     call put('h', h(1,1,:), nCells)
  case default
     stop 'unknown mode'
  end select
  close(uout)

contains

  subroutine put(name, arr, n)
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    real(kind=RKIND), intent(in) :: arr(n)
    character(len=32) :: nm
    nm = name
    write(uout) nm, int(n, 8), arr
  end subroutine put

  subroutine put1(name, x)
    character(len=*), intent(in) :: name
    real(kind=RKIND), intent(in) :: x
    real(kind=RKIND) :: a(1)
    a(1) = x
    call put(name, a, 1)
  end subroutine put1

  subroutine dump_state()
    call put('h', h(1,1,:), nCells)
    call put('u', u(1,1,:), nEdges)
  end subroutine dump_state

  subroutine dump_ic()
    call dump_state()
    call put('h_s', h_s, nCells)
    call put('fVertex', fVertex, nVertices)
    call put('fEdge', fEdge, nEdges)
    call put('dcEdge', dcEdge, nEdges)
    call put('dvEdge', dvEdge, nEdges)
    call put('areaCell', areaCell, nCells)
    call put('areaTriangle', areaTriangle, nVertices)
    call put('kiteAreasOnVertex', reshape(kiteAreasOnVertex, (/3*nVertices/)), 3*nVertices)
    call put('lonCell', lonCell, nCells)
    call put('xCell', xCell, nCells)
    call put('xEdge', xEdge, nEdges)
    call put('coeffs_reconstruct', reshape(coeffs_reconstruct, (/3*maxEdges*nCells/)), 3*maxEdges*nCells)
    call put('edgeNormalVectors', reshape(edgeNormalVectors, (/3*nEdges/)), 3*nEdges)
    call put('cellTangentPlane', reshape(cellTangentPlane, (/6*nCells/)), 6*nCells)
    call put('localVerticalUnitVectors', reshape(localVerticalUnitVectors, (/3*nCells/)), 3*nCells)
    call dump_recon()
  end subroutine dump_ic

  subroutine dump_recon()
    call put('uReconstructX', uReconstructX(1,:), nCells)
    call put('uReconstructY', uReconstructY(1,:), nCells)
    call put('uReconstructZ', uReconstructZ(1,:), nCells)
    call put('uReconstructZonal', uReconstructZonal(1,:), nCells)
    call put('uReconstructMeridional', uReconstructMeridional(1,:), nCells)
  end subroutine dump_recon

  subroutine dump_diag()
    call put('v', v(1,:), nEdges)
    call put('h_edge', h_edge(1,:), nEdges)
    call put('h_vertex', h_vertex(1,:), nVertices)
    call put('circulation', circulation(1,:), nVertices)
    call put('vorticity', vorticity(1,:), nVertices)
    call put('divergence', divergence(1,:), nCells)
    call put('ke', ke(1,:), nCells)
    call put('pv_edge', pv_edge(1,:), nEdges)
    call put('pv_vertex', pv_vertex(1,:), nVertices)
    call put('pv_cell', pv_cell(1,:), nCells)
    call put('vorticity_cell', vorticity_cell(1,:), nCells)
    call put('gradPVn', gradPVn(1,:), nEdges)
    call put('gradPVt', gradPVt(1,:), nEdges)
  end subroutine dump_diag

  subroutine dump_diag_tl()
    call put('v_tl', v_tl(1,:), nEdges)
    call put('h_edge_tl', h_edge_tl(1,:), nEdges)
    call put('h_vertex_tl', h_vertex_tl(1,:), nVertices)
    call put('vorticity_tl', vorticity_tl(1,:), nVertices)
    call put('divergence_tl', divergence_tl(1,:), nCells)
    call put('ke_tl', ke_tl(1,:), nCells)
    call put('pv_edge_tl', pv_edge_tl(1,:), nEdges)
    call put('pv_vertex_tl', pv_vertex_tl(1,:), nVertices)
    call put('pv_cell_tl', pv_cell_tl(1,:), nCells)
    call put('gradPVn_tl', gradPVn_tl(1,:), nEdges)
    call put('gradPVt_tl', gradPVt_tl(1,:), nEdges)
  end subroutine dump_diag_tl

end program msw_ref
