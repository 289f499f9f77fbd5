!  Drop-in adapters for THERMACOND's job-1 hot path over the alt3 C ABI.
!
!  A program unit that uses this module (kappa.f90, after its own `use` lines) gets module
!  procedures with the names and arguments of the reference subroutines it calls, which then
!  resolve here instead of to sub.f90 / cg.f90:
!    calculate_w3_ibz_split_sq   sub.f90:1050-1233   |V3|^2 of the ksubset block (kept on the device)
!    calculate_FGR_mpi2          sub.f90:7199-7499   tetrahedron weights P1sm/P2sm (device)
!    calculate_tauRTA_mpi2       sub.f90:7953-8194   Qvaluempi, Qvalue_Nmpi, Qvalue_Umpi
!    CollisionMatrix_mpi         sub.f90:3034-3210   Collision_Matrixmpi (ksub*ndyn*3, nibz*ndyn*3)
!    cg_quadratic                cg.f90:19-69        on the caller's host matrix and right-hand side
!    update_mpi                  sub.f90:2295-2449   F2mpi of job 3
!  alt3_open, called once after make_kp_reg_tet (kappa.f90:360) and before the first of them,
!  creates the device context from the module state that read_input_fit, read_lattice,
!  read_born and read_fc234 leave (rest.f90:85-443, :3580-3600) and the mesh of module kpoints.
!  The eigen data are the caller's (module eigen, after get_frequencies and subst_eivecs): the
!  |V3|^2 adapter loads the eivec/eival it is passed, with veloc, into the context.  Everything
!  else of kappa.f90 (distribution, Qandtau, mode_kappa, cal_RHS_IBZ, mode_kappa_d,
!  write_kappa, the MPI gathers) stays the reference's own code and reads the module arrays
!  these adapters fill.
!
!  Compiled against the reference's module interfaces (.mod files of modules.f90) and the
!  bind(C) interfaces of alt3_iface.f90.
module alt3_thermacond
  use iso_c_binding
  use alt3_iface
  implicit none
  private
  public :: alt3_open, alt3_close, alt3_h, alt3_precision
  public :: calculate_w3_ibz_split_sq, calculate_FGR_mpi2, calculate_tauRTA_mpi2, CollisionMatrix_mpi, &
            cg_quadratic, update_mpi

  type(c_ptr), save :: alt3_h = c_null_ptr
  integer(c_int), save :: alt3_precision = ALT3_V3_FP64
  ! the temperature of the last calculate_tauRTA_mpi2: CollisionMatrix_mpi and update_mpi of the
  ! reference read the occupations distribution() left for it from module exactBTE2
  real(c_double), save :: tempk_last = -1

  ! copies handed to alt3_create (it copies them in turn; kept until alt3_close)
  real(c_double), allocatable, target, save :: s_mass(:), s_atompos(:, :), s_fc2(:), s_fc3(:), s_zeu(:)
  real(c_double), target, save :: s_r123(3, 3), s_latp(6), s_prim(9), s_epsil(9)
  integer(c_int), allocatable, target, save :: s_iat0(:), s_fc2a(:, :), s_fc2x(:, :), s_fc3a(:, :), s_fc3x(:, :)

contains

  subroutine alt3_open(precision)
    use lattice, only: r01, r02, r03, latticeparameters, primitivelattice
    use params, only: classical
    use atoms_force_constants, only: natoms0, natoms, atompos, iatomcell0, atom0
    use svd_stuff, only: nterms, iatomterm_2, ixyzterm_2, fcs_2, iatomterm_3, ixyzterm_3, fcs_3
    use kpoints, only: nc
    use born, only: born_flag, epsil
    integer, intent(in), optional :: precision
    type(alt3_system) :: s
    integer :: i, j, k
    if (present(precision)) alt3_precision = precision
    allocate(s_mass(natoms0), s_atompos(3, natoms), s_iat0(natoms))
    do i = 1, natoms0
       s_mass(i) = atom0(i)%mass
    end do
    s_r123(:, 1) = [r01%x, r01%y, r01%z]
    s_r123(:, 2) = [r02%x, r02%y, r02%z]
    s_r123(:, 3) = [r03%x, r03%y, r03%z]
    s_latp = latticeparameters                       ! already scaled by read_input_fit
    s_prim = reshape(primitivelattice, [9])          ! column j = primitive vector j
    s_atompos = atompos(:, 1:natoms)
    s_iat0 = iatomcell0(1:natoms)
    allocate(s_fc2a(2, nterms(2)), s_fc2x(2, nterms(2)), s_fc2(nterms(2)))
    s_fc2a = iatomterm_2(:, 1:nterms(2))
    s_fc2x = ixyzterm_2(:, 1:nterms(2))
    s_fc2 = fcs_2(1:nterms(2))
    allocate(s_fc3a(3, nterms(3)), s_fc3x(3, nterms(3)), s_fc3(nterms(3)))
    s_fc3a = iatomterm_3(:, 1:nterms(3))
    s_fc3x = ixyzterm_3(:, 1:nterms(3))
    s_fc3 = fcs_3(1:nterms(3))
    ! read_born: epsil(i,j) and atom0(k)%charge(i,j), row-major for the C side
    allocate(s_zeu(9 * natoms0))
    s_epsil = 0
    s_zeu = 0
    if (born_flag /= 0) then
       do i = 1, 3
          do j = 1, 3
             s_epsil(3 * (i - 1) + j) = epsil(i, j)
             do k = 1, natoms0
                s_zeu(9 * (k - 1) + 3 * (i - 1) + j) = atom0(k)%charge(i, j)
             end do
          end do
       end do
    end if
    s%natoms0 = natoms0
    s%mass = c_loc(s_mass)
    s%r123 = c_loc(s_r123)
    s%latparams = c_loc(s_latp)
    s%primlatt = c_loc(s_prim)
    s%natoms = natoms
    s%atompos = c_loc(s_atompos)
    s%iatomcell0 = c_loc(s_iat0)
    s%nt2 = nterms(2)
    s%fc2_atoms = c_loc(s_fc2a)
    s%fc2_xyz = c_loc(s_fc2x)
    s%fc2 = c_loc(s_fc2)
    s%nt3 = nterms(3)
    s%fc3_atoms = c_loc(s_fc3a)
    s%fc3_xyz = c_loc(s_fc3x)
    s%fc3 = c_loc(s_fc3)
    s%nc = nc
    s%classical = classical
    s%host_only = 0
    s%born_flag = born_flag
    s%epsil = c_loc(s_epsil)
    s%zeu = c_loc(s_zeu)
    call alt3_check(alt3_create(s, alt3_h), 'alt3_create')
  end subroutine alt3_open

  subroutine alt3_close()
    if (c_associated(alt3_h)) call alt3_check(alt3_destroy(alt3_h), 'alt3_destroy')
    alt3_h = c_null_ptr
    if (allocated(s_mass)) deallocate(s_mass, s_atompos, s_iat0, s_fc2a, s_fc2x, s_fc2, s_fc3a, s_fc3x, s_fc3, s_zeu)
  end subroutine alt3_close

  ! sub.f90:1050.  The block's |V3|^2 stays on the device (v33s8 of module phi3 is not
  ! filled: only the split-file path of job 4 reads it on the host).
  subroutine calculate_w3_ibz_split_sq(ibz_subset, nv3_split, ndn, nk, kp, eival, eivec, nv3s)
    use eigen, only: veloc
    integer :: ibz_subset(2), nv3_split, ndn, nk, nv3s
    real(8) :: kp(3, nk)
    real(8), target :: eival(ndn, nk)
    complex(8), target :: eivec(ndn, ndn, nk)
    real(c_double), allocatable, target :: vel(:, :, :)
    allocate(vel(3, ndn, nk))
    vel = veloc(:, :, 1:nk)
    call alt3_check(alt3_set_eig(alt3_h, c_loc(eival), c_loc(eivec), c_loc(vel)), 'alt3_set_eig')
    call alt3_check(alt3_v3sq_ibz(alt3_h, ibz_subset(1) - 1, ibz_subset(2) - 1, alt3_precision, c_null_ptr), &
         'calculate_w3_ibz_split_sq')
    nv3s = (ibz_subset(2) - ibz_subset(1) + 1) * nk * ndn**3
  end subroutine calculate_w3_ibz_split_sq

  ! sub.f90:7199.  The module arrays the reference routine sets up on its way (allocate_iter,
  ! cal_freq: frequency and dist) are set up the same way; P1smpi/P2smpi stay on the device.
  subroutine calculate_FGR_mpi2(ksubset, ndn, nv3o)
    use kpoints, only: nibz, nkc
    use exactBTE2, only: allocate_iter
    integer :: ksubset(2), ndn, nv3o
    call allocate_iter(nibz, nkc, ndn)
    call cal_freq(nkc, ndn)
    call alt3_check(alt3_fgr(alt3_h, c_null_ptr, c_null_ptr), 'calculate_FGR_mpi2')
  end subroutine calculate_FGR_mpi2

  ! sub.f90:7953: Q, Q_N and Q_U of the block into Qvaluempi, Qvalue_Nmpi, Qvalue_Umpi (ksub, ndn)
  subroutine calculate_tauRTA_mpi2(ibz_subset, ndn, tempk)
    use exactBTE2, only: Qvaluempi, Qvalue_Nmpi, Qvalue_Umpi
    integer :: ibz_subset(2), ndn
    real(8) :: tempk
    real(c_double), allocatable, target :: q(:, :), qn(:, :), qu(:, :)
    integer :: ksub
    ksub = ibz_subset(2) - ibz_subset(1) + 1
    allocate(q(ksub, ndn), qn(ksub, ndn), qu(ksub, ndn))
    call alt3_check(alt3_tau(alt3_h, tempk, c_loc(q), c_loc(qn), c_loc(qu), c_null_ptr, c_null_ptr, c_null_ptr), &
         'calculate_tauRTA_mpi2')
    Qvaluempi(1:ksub, 1:ndn) = q
    Qvalue_Nmpi(1:ksub, 1:ndn) = qn
    Qvalue_Umpi(1:ksub, 1:ndn) = qu
    tempk_last = tempk
  end subroutine calculate_tauRTA_mpi2

  ! sub.f90:3034: the rows of the block into Collision_Matrixmpi
  subroutine CollisionMatrix_mpi(ksbst, ndn)
    use exactBTE2, only: Collision_Matrixmpi
    integer :: ksbst(2), ndn
    real(c_double), allocatable, target :: cm(:, :)
    if (tempk_last < 0) call alt3_fail('CollisionMatrix_mpi before calculate_tauRTA_mpi2')
    allocate(cm(size(Collision_Matrixmpi, 1), size(Collision_Matrixmpi, 2)))
    call alt3_check(alt3_collision(alt3_h, tempk_last, c_loc(cm), c_null_ptr), 'CollisionMatrix_mpi')
    Collision_Matrixmpi = cm
  end subroutine CollisionMatrix_mpi

  ! cg.f90:19: the reference's arguments (host matrix and right-hand side)
  subroutine cg_quadratic(n, a, b, vmin, funcmin, max_iter, n_dig_accur)
    integer :: n, max_iter, n_dig_accur
    real(8), target :: a(n, n), b(n), vmin(n), funcmin
    integer(c_int), target :: iters
    call alt3_check(alt3_cg(alt3_h, n, c_loc(a), c_loc(b), c_loc(vmin), max_iter, n_dig_accur, c_loc(funcmin), &
         c_loc(iters)), 'cg_quadratic')
  end subroutine cg_quadratic

  ! sub.f90:2295 (job 3): F2mpi = F_RTA + A(FFF)/Qvalue for the block
  subroutine update_mpi(ndn, ibz_subset)
    use exactBTE2, only: FFF, F_RTA, F2mpi
    integer :: ndn, ibz_subset(2)
    real(c_double), allocatable, target :: f(:, :, :), frta(:, :, :), f2(:, :, :)
    if (tempk_last < 0) call alt3_fail('update_mpi before calculate_tauRTA_mpi2')
    allocate(f, source=FFF)
    allocate(frta, source=F_RTA)
    allocate(f2(size(F2mpi, 1), size(F2mpi, 2), size(F2mpi, 3)))
    call alt3_check(alt3_update(alt3_h, tempk_last, c_loc(f), c_loc(frta), c_loc(f2)), 'update_mpi')
    F2mpi = f2
  end subroutine update_mpi

  subroutine alt3_fail(msg)
    character(*), intent(in) :: msg
    write(*, *) 'alt3_thermacond: ', msg
    error stop 3
  end subroutine alt3_fail

end module alt3_thermacond
