!  THERMACOND job 1 (kappa.f90:178-360 and :389-871, one process, ksubset = [1, nibz]) with
!  the reference's own routines for everything but the hot path, which the adapters of module
!  alt3_thermacond replace: the same `call` statements as kappa.f90, resolved to the adapters
!  because this program unit uses the module.  Input readers (read_params, read_input_fit,
!  read_lattice, read_born, read_fc234), the mesh and symmetry setup, get_frequencies and
!  subst_eivecs (CPU, the reference's eigen solver), distribution, Qandtau, mode_kappa,
!  calculate_RTA_distributionfunction, cal_RHS_IBZ, mode_kappa_d and write_kappa are the
!  reference's; calculate_w3_ibz_split_sq, calculate_FGR_mpi2, calculate_tauRTA_mpi2,
!  CollisionMatrix_mpi and cg_quadratic run on the GPU through libalt3.
!
!  Writes kappa_RTA.dat and tau_RTA.dat (kappa.f90:560-678), kappa_direct-cg.dat and
!  tau_RTAandeff-cg.dat (kappa.f90:701-869) at T = tmin, with the reference's write_kappa.
!  Built only where the reference sources are (their objects are linked in).
!
!  usage: alt3_job1_ref [fp64 | mixed]
program alt3_job1_ref
  use ios
  use kpoints
  use om_dos
  use atoms_force_constants
  use phi3
  use phi3_sy
  use eigen
  use params
  use constants
  use born
  use exactBTE2
  use tetrahedron
  use lattice
  use svd_stuff
  use alt3_iface, only: ALT3_V3_FP64, ALT3_V3_MIXED
  use alt3_thermacond
  implicit none
  integer :: ksubset(2), ksub, nv3_split, nv3o, s, is, js, ks, prec
  real(8) :: tempk, fm
  real(8), allocatable :: dkc(:)
  character(len=16) :: arg

  prec = ALT3_V3_FP64
  if (command_argument_count() > 0) then
     call get_command_argument(1, arg)
     if (trim(arg) == 'mixed') prec = ALT3_V3_MIXED
  end if

  ! the preamble of kappa.f90 (inputs, mesh, irreducible zone, eigen solve, tetrahedra)
  call read_params
  call read_input_fit
  call read_lattice
  call read_born
  nkc = nc(1) * nc(2) * nc(3)
  call read_fc234
  ndyn = 3 * natoms0
  allocate(kpc(3, nkc), mappos(nkc), wk(nkc))
  call make_kp_reg(nc, g1, g2, g3, shft, kpc, wk)
  call get_negatives(nkc, kpc, mappos, npos)
  call get_weights(nkc, kpc)
  call allocate_eig(ndyn, nkc, nkc)
  allocate(dkc(nkc))
  dkc = 0
  call get_frequencies(nkc, kpc, dkc, ndyn, eigenval, ndyn, eigenvec, ucors, veloc)
  call subst_eivecs(ndyn, nkc, eigenvec, kpc, npos, mappos)
  call allocate_tetra(nc(1), nc(2), nc(3), wmesh, ndyn)
  call make_kp_reg_tet(nc(1), nc(2), nc(3), shft(1), shft(2), shft(3), kpc, wk)

  call alt3_open(prec)

  ! job 1 on one "rank": the whole IBZ
  ksubset(1) = 1
  ksubset(2) = nibz
  ksub = nibz
  nv3_split = nkc * ksub * ndyn**3
  call calculate_w3_ibz_split_sq(ksubset, nv3_split, ndyn, nkc, kpc, eigenval, eigenvec, nv3o)
  call calculate_FGR_mpi2(ksubset, ndyn, nv3o)
  allocate(FFF(nibz, ndyn, 3), Qvaluempi(ksub, ndyn), Qvalue_Nmpi(ksub, ndyn), Qvalue_Umpi(ksub, ndyn))
  allocate(Collision_Matrixmpi(ksub * ndyn * 3, nibz * ndyn * 3), RHS(nibz * ndyn * 3), DirectSolution(nibz * ndyn * 3))

  tempk = tmin
  call distribution(nkc, ndyn, tempk)
  call calculate_tauRTA_mpi2(ksubset, ndyn, tempk)
  Qvalue = Qvaluempi                     ! the gathers of kappa.f90:641-643, one rank
  Qvalue_N = Qvalue_Nmpi
  Qvalue_U = Qvalue_Umpi
  call Qandtau(ndyn, tempk)
  call calculate_RTA_distributionfunction(nibz, nkc, kibz, ndyn, tempk, veloc, eigenval)
  call mode_kappa(ndyn, tempk, veloc, eigenval)
  open(3450, file='tau_RTA.dat', status='unknown')
  open(6450, file='kappa_RTA.dat', status='unknown')
  call write_kappa(ndyn, tempk, veloc, eigenval, 6450, 3450)
  close(3450)
  close(6450)

  allocate(Collision_Matrix(nibz * ndyn * 3, nibz * ndyn * 3))
  call CollisionMatrix_mpi(ksubset, ndyn)
  Collision_Matrix = Collision_Matrixmpi
  call cal_RHS_IBZ(ndyn, nkc, tempk, veloc, eigenval)
  s = 0
  do is = 1, nibz
     do js = 1, ndyn
        do ks = 1, 3
           s = s + 1
           DirectSolution(s) = FFF(is, js, ks)
        end do
     end do
  end do
  call cg_quadratic(nibz * ndyn * 3, Collision_Matrix, RHS, DirectSolution, fm, 100, 2)
  call mode_kappa_d(ndyn, tempk, veloc, eigenval, DirectSolution)
  open(7450, file='tau_RTAandeff-cg.dat', status='unknown')
  open(8450, file='kappa_direct-cg.dat', status='unknown')
  call write_kappa(ndyn, tempk, veloc, eigenval, 8450, 7450)
  close(7450)
  close(8450)

  call alt3_close()
  write(*, '(a)') 'alt3_job1_ref: done'
end program alt3_job1_ref
