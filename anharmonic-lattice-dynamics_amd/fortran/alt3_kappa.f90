!  THERMACOND jobs 1 and 3 (kappa.f90:283-851, :1046-1487) driven through the alt3 C ABI.
!
!  Reads the reference's input files from the working directory (latdyn.params,
!  structure.params, lat_fc.dat, fc2.dat, fc3.dat; readers rest.f90:2-443), runs the
!  three-phonon path on the GPU stage by stage in the order of kappa.f90, and writes the
!  reference's outputs with write_kappa's formats (sub.f90:1818-1916):
!    tau_RTA.dat, kappa_RTA.dat                 (RTA,      kappa.f90:562-678, :1190-1304)
!    job 1: tau_RTAandeff-cg.dat, kappa_direct-cg.dat   (CG / BTE, kappa.f90:703-851)
!    job 3: tau_RTAandeff_iter.dat, kappa_iter.dat       (iterative BTE, kappa.f90:1373-1487)
!    job 4: CM-1D.dat, RHS_IBZ.dat, DS.dat (the CG system, kappa.f90:2577-2625) after the RTA
!    job 5: reads those three, CG and kappa_BTE as job 1 (kappa.f90:2635-2703)
!  The job is the one latdyn.params names (3, 4, 5, else 1); the temperature is tmin
!  (kappa.f90:565).  iso = 1 in latdyn.params adds isotope scattering (cal_giso, P_iso_split,
!  calculate_tauRTA_iso, CollisionMatrix_iso, update_iso; kappa.f90:522-543, :574-578,
!  :710-716, :1401-1407).
!
!  usage: alt3_kappa [fp64 | mixed | fp64ref] [--host-only] [--eig FILE]
!    --host-only builds the mesh and symmetry tables only (no GPU) and prints their sizes.
!    --eig FILE  takes the eigen data from FILE instead of the device eigen solve (alt3_set_eig):
!                one stream of eigenval(nb,nk) real(8), eigenvec(nb,nb,nk) complex(8) and
!                veloc(3,nb,nk) real(8), the arrays get_frequencies + subst_eivecs leave in
!                module eigen (kappa.f90:283-300), so that the rest of the path and the output
!                files can be set beside a reference run on the same eigenvectors.
program alt3_kappa
  use iso_c_binding
  use alt3_iface
  implicit none

  real(c_double), parameter :: k_b = 1.3806504d-23, h_plank = 6.62606896d-34, c_light = 2.99792458d+8
  real(4), parameter :: cnst = 521.1098918      ! REAL*4, modules.f90:18

  ! latdyn.params
  integer :: nc(3), wmesh, ntemp, job, ksub_size, iso, classical
  logical :: dos_bs_only, verbose
  real(8) :: shft(3), normald(3), wmax, width, etaz, tmin, tmax
  ! structure.params
  real(c_double), target :: latparams(6), primlatt(9)
  real(8) :: scale
  integer :: natom_type, natoms0
  real(8), allocatable :: mas(:)
  character(len=2), allocatable :: atname(:)
  real(c_double), allocatable, target :: giso(:)
  integer, allocatable :: atom_type(:)
  real(c_double), allocatable, target :: mass(:)
  ! lat_fc.dat
  real(c_double), target :: r123(3, 3)
  integer :: nterms(4), nsmax, natoms
  real(c_double), allocatable, target :: atompos(:, :)
  integer(c_int), allocatable, target :: iatomcell0(:)
  ! fc2.dat / fc3.dat
  integer(c_int), allocatable, target :: fc2_atoms(:, :), fc2_xyz(:, :), fc3_atoms(:, :), fc3_xyz(:, :)
  real(c_double), allocatable, target :: fc2(:), fc3(:)
  ! dielectric.params (read_born): epsil(i,j) and charge(i,j) of each atom, row-major for the C side
  integer :: born_flag
  real(c_double), target :: epsil_rm(9)
  real(c_double), allocatable, target :: zeu_rm(:)

  type(alt3_system) :: sys
  type(c_ptr) :: ctx
  integer(c_int) :: nk, nibz, nb, nops, prec, n
  integer(c_int), target :: iters
  real(c_double), allocatable, target :: eigenval(:, :), veloc(:, :, :), kibz(:, :)
  integer(c_int), allocatable, target :: mapinv(:)
  real(c_double), allocatable, target :: q(:, :), qn(:, :), qu(:, :), taun(:, :), tauu(:, :), tautot(:, :)
  real(c_double), allocatable, target :: frta(:, :, :), taueff(:, :), x(:)
  real(c_double), allocatable, target :: kappa(:), khist(:), tehist(:), cm(:, :), rhs(:)
  real(c_double), target :: fmin
  real(c_double) :: tempk
  character(len=32) :: arg
  character(len=4096) :: eig_file
  complex(c_double_complex), allocatable, target :: eigenvec(:, :, :)
  logical :: host_only
  integer :: i, l, a, ia, it, u

  prec = ALT3_V3_FP64
  host_only = .false.
  eig_file = ''
  i = 0
  do while (i < command_argument_count())
     i = i + 1
     call get_command_argument(i, arg)
     select case (trim(arg))
     case ('fp64'); prec = ALT3_V3_FP64
     case ('mixed'); prec = ALT3_V3_MIXED
     case ('fp64ref'); prec = ALT3_V3_FP64_REF
     case ('--host-only'); host_only = .true.
     case ('--eig')
        i = i + 1
        call get_command_argument(i, eig_file)
        if (len_trim(eig_file) == 0) then
           write(*, *) 'alt3_kappa: --eig needs a file'
           stop 2
        end if
     case default
        write(*, *) 'alt3_kappa: unknown argument ', trim(arg)
        stop 2
     end select
  end do

  call read_latdyn()
  call read_structure()
  call read_lat_fc()
  call read_fc('fc2.dat', 2, nterms(2), fc2_atoms, fc2_xyz, fc2)
  call read_fc('fc3.dat', 3, nterms(3), fc3_atoms, fc3_xyz, fc3)
  call read_born()

  ! the module state of read_lattice/read_fc234, handed over as one struct
  latparams(1:3) = latparams(1:3) * scale          ! rest.f90:92
  sys%natoms0 = natoms0
  sys%mass = c_loc(mass)
  sys%r123 = c_loc(r123)
  sys%latparams = c_loc(latparams)
  sys%primlatt = c_loc(primlatt)
  sys%natoms = natoms
  sys%atompos = c_loc(atompos)
  sys%iatomcell0 = c_loc(iatomcell0)
  sys%nt2 = nterms(2)
  sys%fc2_atoms = c_loc(fc2_atoms)
  sys%fc2_xyz = c_loc(fc2_xyz)
  sys%fc2 = c_loc(fc2)
  sys%nt3 = nterms(3)
  sys%fc3_atoms = c_loc(fc3_atoms)
  sys%fc3_xyz = c_loc(fc3_xyz)
  sys%fc3 = c_loc(fc3)
  sys%nc = nc
  sys%classical = classical
  sys%host_only = merge(1, 0, host_only)
  sys%born_flag = born_flag
  sys%epsil = c_loc(epsil_rm)
  sys%zeu = c_loc(zeu_rm)
  call alt3_check(alt3_create(sys, ctx), 'alt3_create')
  call alt3_check(alt3_sizes(ctx, nk, nibz, nb, nops), 'alt3_sizes')
  write(*, '(a,i0,a,i0,a,i0,a,i0)') 'alt3_kappa: nk=', nk, ' nibz=', nibz, ' nb=', nb, ' nops=', nops
  if (host_only) then
     call alt3_check(alt3_destroy(ctx), 'alt3_destroy')
     stop
  end if

  allocate(eigenval(nb, nk), veloc(3, nb, nk), kibz(3, nibz), mapinv(nibz))
  allocate(q(nibz, nb), qn(nibz, nb), qu(nibz, nb), taun(nibz, nb), tauu(nibz, nb), tautot(nibz, nb))
  allocate(frta(nibz, nb, 3), taueff(nibz, nb), kappa(nb * 9))
  n = nibz * nb * 3
  allocate(x(n))

  if (len_trim(eig_file) > 0) then
     ! eigen data of another solver (e.g. the reference's get_frequencies + subst_eivecs)
     allocate(eigenvec(nb, nb, nk))
     open(newunit=u, file=trim(eig_file), access='stream', form='unformatted', status='old')
     read(u) eigenval, eigenvec, veloc
     close(u)
     call alt3_check(alt3_set_eig(ctx, c_loc(eigenval), c_loc(eigenvec), c_loc(veloc)), 'alt3_set_eig')
     deallocate(eigenvec)
  else
     ! get_frequencies (kappa.f90:283); eigenvectors stay on the device
     call alt3_check(alt3_eig(ctx, c_loc(eigenval), c_null_ptr, c_loc(veloc)), 'alt3_eig')
  end if
  call alt3_check(alt3_get_mesh(ctx, c_null_ptr, c_null_ptr, c_null_ptr, c_loc(mapinv), c_loc(kibz), &
       c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr), 'alt3_get_mesh')
  tempk = tmin                                     ! kappa.f90:565, :2649
  if (iso == 1) then
     ! cal_giso and P_iso_split (kappa.f90:522-543): g of each atom from its element, the
     ! isotope term on the device; tau, collision and update then take their iso forms
     allocate(giso(natoms0))
     do i = 1, natoms0
        call alt3_check(alt3_isotope_g(trim(atname(atom_type(i)))//c_null_char, giso(i)), 'alt3_isotope_g')
     end do
     call alt3_check(alt3_set_iso(ctx, c_loc(giso), c_null_ptr), 'alt3_set_iso')
     write(*, '(a,99es12.4)') 'alt3_kappa: iso = 1, g per atom ', giso
  end if

  if (job == 5) then
     ! read_CM_1D_mpi, read_RHS, read_DS (sub.f90:5401-5470), then cg_quadratic and
     ! mode_kappa_d; the scattering rates of the tau file are not known here, as in the reference
     allocate(cm(n, n), rhs(n))
     call read_rows('CM-1D.dat', n, cm)
     call read_column('RHS_IBZ.dat', n, rhs)
     call read_column('DS.dat', n, x)
     call alt3_check(alt3_cg(ctx, n, c_loc(cm), c_loc(rhs), c_loc(x), 100, 2, c_loc(fmin), c_loc(iters)), 'alt3_cg')
     taueff = 0
     taun = 0
     tauu = 0
     tautot = 0
     call alt3_check(alt3_kappa_bte(ctx, tempk, c_loc(x), c_loc(kappa), c_loc(taueff)), 'alt3_kappa_bte')
     open(10, file='kappa_direct-cg.dat', status='unknown')
     open(11, file='tau_RTAandeff-cg.dat', status='unknown')
     call write_kappa(10, 11)
     close(10)
     close(11)
     write(*, '(a,i0,a,es14.6)') 'alt3_kappa: job 5 cg iterations ', iters, ', funcmin ', fmin
     call alt3_check(alt3_destroy(ctx), 'alt3_destroy')
     stop
  end if
  ! calculate_w3_ibz_split_sq over the whole IBZ (kappa.f90:440), calculate_FGR_mpi2 (:485)
  call alt3_check(alt3_v3sq_ibz(ctx, 0, nibz - 1, prec, c_null_ptr), 'alt3_v3sq_ibz')
  call alt3_check(alt3_fgr(ctx, c_null_ptr, c_null_ptr), 'alt3_fgr')

  ! distribution + calculate_tauRTA_mpi2 + Qandtau (kappa.f90:567-667)
  call alt3_check(alt3_tau(ctx, tempk, c_loc(q), c_loc(qn), c_loc(qu), c_loc(taun), c_loc(tauu), c_loc(tautot)), &
       'alt3_tau')
  ! calculate_RTA_distributionfunction + mode_kappa (kappa.f90:672-677)
  taueff = 0
  call alt3_check(alt3_kappa_rta(ctx, tempk, c_loc(kappa), c_loc(frta), c_loc(taueff)), 'alt3_kappa_rta')
  open(10, file='kappa_RTA.dat', status='unknown')
  open(11, file='tau_RTA.dat', status='unknown')
  call write_kappa(10, 11)
  close(10)
  close(11)

  if (job == 3) then
     ! iterative BTE (kappa.f90:1373-1487): max_iter = 150, update_mix = 0.4 and the stop
     ! normk < 0.0001, the last two REAL*4 literals (rest.f90:36,44)
     allocate(khist(150 * nb * 9), tehist(150))
     call alt3_check(alt3_kappa_iter(ctx, tempk, 150, real(0.4, c_double), real(0.0001, c_double), c_loc(kappa), &
          c_loc(khist), c_loc(tehist), c_null_ptr, c_loc(iters)), 'alt3_kappa_iter')
     open(10, file='kappa_iter.dat', status='unknown')
     open(11, file='tau_RTAandeff_iter.dat', status='unknown')
     do it = 1, iters                  ! (write_kappa loops over the host variable i)
        ! mode_kappa of iteration it: kappa, and tauinv_eff cleared but for its last mode
        kappa = khist((it - 1) * nb * 9 + 1:it * nb * 9)
        taueff = 0
        taueff(nibz, nb) = tehist(it)
        write(10, *) '================= iteration', it
        call write_kappa(10, 11)
     end do
     close(10)
     close(11)
     write(*, '(a,i0)') 'alt3_kappa: job 3 iterations ', iters
     call alt3_check(alt3_destroy(ctx), 'alt3_destroy')
     stop
  end if

  if (job == 4) then
     ! the collision system for a later job 5: CM-1D.dat row by row, RHS_IBZ.dat, and DS.dat =
     ! F_RTA flattened (i, l, alpha), alpha fastest (kappa.f90:2577-2625).  One value per line as
     ! there, but with 17 significant digits (the reference's list-directed output keeps fewer),
     ! so that job 5's list-directed reads (read_CM_1D_mpi, read_RHS, read_DS) get every bit back
     allocate(cm(n, n), rhs(n))
     call alt3_check(alt3_collision(ctx, tempk, c_loc(cm), c_loc(rhs)), 'alt3_collision')
     open(10, file='CM-1D.dat', status='unknown')
     do i = 1, n
        do ia = 1, n
           write(10, '(es25.17e3)') cm(i, ia)
        end do
     end do
     close(10)
     open(10, file='RHS_IBZ.dat', status='unknown')
     do i = 1, n
        write(10, '(es25.17e3)') rhs(i)
     end do
     close(10)
     open(10, file='DS.dat', status='unknown')
     do i = 1, nibz
        do l = 1, nb
           do a = 1, 3
              write(10, '(es25.17e3)') frta(i, l, a)
           end do
        end do
     end do
     close(10)
     write(*, '(a,i0)') 'alt3_kappa: job 4 wrote the collision system, n = ', n
     call alt3_check(alt3_destroy(ctx), 'alt3_destroy')
     stop
  end if

  ! CollisionMatrix_mpi + cal_RHS_IBZ (kappa.f90:719-822); the matrix stays on the device
  call alt3_check(alt3_collision(ctx, tempk, c_null_ptr, c_null_ptr), 'alt3_collision')
  ! DirectSolution = F_RTA flattened (i, l, alpha), alpha fastest (kappa.f90:835-843)
  do i = 1, nibz
     do l = 1, nb
        do a = 1, 3
           x(3 * ((i - 1) * nb + l - 1) + a) = frta(i, l, a)
        end do
     end do
  end do
  ! cg_quadratic(nibz*ndyn*3, Collision_Matrix, RHS, DirectSolution, fm, 100, 2) (kappa.f90:846)
  call alt3_check(alt3_cg(ctx, n, c_null_ptr, c_null_ptr, c_loc(x), 100, 2, c_loc(fmin), c_loc(iters)), 'alt3_cg')
  ! mode_kappa_d (kappa.f90:851), tauinv_eff accumulating on the RTA values
  call alt3_check(alt3_kappa_bte(ctx, tempk, c_loc(x), c_loc(kappa), c_loc(taueff)), 'alt3_kappa_bte')
  open(10, file='kappa_direct-cg.dat', status='unknown')
  open(11, file='tau_RTAandeff-cg.dat', status='unknown')
  call write_kappa(10, 11)
  close(10)
  close(11)
  write(*, '(a,i0,a,es14.6)') 'alt3_kappa: cg iterations ', iters, ', funcmin ', fmin

  call alt3_check(alt3_destroy(ctx), 'alt3_destroy')

contains

  subroutine read_latdyn()
    ! read_params, rest.f90:13-51
    character(len=256) :: v3path
    integer :: u
    open(newunit=u, file='latdyn.params', status='old')
    read(u, *) nc, dos_bs_only
    read(u, *) shft
    read(u, *) normald
    read(u, *) wmesh, wmax
    read(u, *) width, etaz
    read(u, *) verbose
    read(u, *) tmin, tmax, ntemp
    read(u, *) job
    read(u, *) ksub_size
    read(u, '(a)') v3path
    read(u, *) iso
    read(u, *) classical
    close(u)
  end subroutine read_latdyn

  subroutine read_structure()
    ! read_input_fit, rest.f90:87-127
    character(len=256) :: junk
    real(8) :: pos(3)
    integer :: u, k, label, ty
    open(newunit=u, file='structure.params', status='old')
    read(u, *) latparams
    read(u, *) primlatt
    read(u, *) scale
    do k = 1, 4
       read(u, '(a)') junk
    end do
    read(u, *) natom_type
    allocate(mas(natom_type))
    read(u, *) mas
    allocate(atname(natom_type))
    read(u, *) atname                            ! element names
    read(u, *) natoms0
    do k = 1, 8
       read(u, '(a)') junk
    end do
    allocate(atom_type(natoms0), mass(natoms0))
    do k = 1, natoms0
       read(u, *) label, ty, pos
       atom_type(label) = ty
    end do
    close(u)
    do k = 1, natoms0
       mass(k) = mas(atom_type(k))
    end do
  end subroutine read_structure

  subroutine read_lat_fc()
    ! read_lattice, rest.f90:166-289 (file written by FOCEX/ios.f90:2278-2333)
    character(len=512) :: line
    real(8) :: pos(3), d
    integer :: u, k, n1, ia, i0, cell(3)
    open(newunit=u, file='lat_fc.dat', status='old')
    do
       read(u, '(a)') line
       if (index(line, 'translation vectors') > 0) exit
    end do
    do k = 1, 3
       read(u, *) r123(:, k)                     ! r123(:,k) = r_k, the C rows
    end do
    read(u, '(a)') line
    read(u, *) n1
    if (n1 /= natoms0) then
       write(*, *) 'natoms0 in lat_fc.dat incompatible with structure.params', n1, natoms0
       stop 1
    end if
    do k = 1, n1
       read(u, '(a)') line
    end do
    do k = 1, 1 + 1 + 8 + 2 + 1                  ! ranks, shells, groups headers and records
       read(u, '(a)') line
    end do
    read(u, *) nterms
    do k = 1, 3
       read(u, '(a)') line
    end do
    read(u, *) nsmax, natoms
    allocate(atompos(3, natoms), iatomcell0(natoms))
    do k = 1, natoms
       read(u, *) ia, pos, i0, cell, d
       atompos(:, ia) = pos
       iatomcell0(ia) = i0
    end do
    close(u)
  end subroutine read_lat_fc

  ! read_born (rest.f90:3580-3600): born_flag, epsil(i,1:3) per record, then charge(i,1:3) of
  ! each atom of the primitive cell, three records per atom; the library adds nonanal's term
  ! to the dynamical matrices when born_flag /= 0
  subroutine read_born()
    integer :: u, k, i, j
    real(c_double) :: e(3, 3), z(3, 3)
    allocate(zeu_rm(9 * natoms0))
    zeu_rm = 0
    open(newunit=u, file='dielectric.params', status='old')
    read(u, *) born_flag
    do i = 1, 3
       read(u, *) (e(i, j), j = 1, 3)
    end do
    do k = 1, natoms0
       do i = 1, 3
          read(u, *) (z(i, j), j = 1, 3)
       end do
       do i = 1, 3
          do j = 1, 3
             zeu_rm(9 * (k - 1) + 3 * (i - 1) + j) = z(i, j)
          end do
       end do
    end do
    close(u)
    do i = 1, 3
       do j = 1, 3
          epsil_rm(3 * (i - 1) + j) = e(i, j)
       end do
    end do
  end subroutine read_born

  subroutine read_fc(fname, rank, nt, atoms, xyz, val)
    ! read_fc234, rest.f90:337-402
    character(*), intent(in) :: fname
    integer, intent(in) :: rank, nt
    integer(c_int), allocatable, intent(out) :: atoms(:, :), xyz(:, :)
    real(c_double), allocatable, intent(out) :: val(:)
    character(len=256) :: line
    integer :: u, k, r, term, group, ia(3), ix(3)
    allocate(atoms(rank, nt), xyz(rank, nt), val(nt))
    open(newunit=u, file=fname, status='old')
    read(u, '(a)') line
    do k = 1, nt
       read(u, *) term, group, (ia(r), ix(r), r = 1, rank), val(k)
       atoms(:, k) = ia(1:rank)
       xyz(:, k) = ix(1:rank)
    end do
    close(u)
  end subroutine read_fc

  subroutine read_column(fname, cnt, v)
    ! one list-directed value per line (read_RHS / read_CM_1D_mpi / read_DS, sub.f90:5367-5470)
    character(*), intent(in) :: fname
    integer, intent(in) :: cnt
    real(c_double), intent(out) :: v(*)
    integer :: u, k
    open(newunit=u, file=fname, status='old')
    do k = 1, cnt
       read(u, *) v(k)
    end do
    close(u)
  end subroutine read_column

  subroutine read_rows(fname, m, a)
    ! CM-1D.dat holds the matrix row by row; read_CM_1D_mpi fills CM3D(i,j) with j fastest
    ! (sub.f90:5424-5431)
    character(*), intent(in) :: fname
    integer, intent(in) :: m
    real(c_double), intent(out) :: a(m, m)
    integer :: u, r, c
    open(newunit=u, file=fname, status='old')
    do r = 1, m
       do c = 1, m
          read(u, *) a(r, c)
       end do
    end do
    close(u)
  end subroutine read_rows

  real(8) function nbe(w, t)
    ! nbe, extratools.f90:257-288
    real(8), intent(in) :: w, t
    real(8) :: z
    z = w / t
    if (classical == 1) then
       nbe = 1 / z
    else if (z > 60) then
       nbe = 0
    else if (z < dble(0.0010)) then
       nbe = 1 / (z * (1 + z / 2 * (1 + z / 3)))
    else
       nbe = 1 / (exp(z) - 1)
    end if
  end function nbe

  subroutine write_kappa(ukap, utau)
    ! write_kappa, sub.f90:1818-1916 (formats 8, 78 and 10 there; G editing of the two
    ! integers of format 10 is written as I14)
    integer, intent(in) :: ukap, utau
    real(8) :: kap(nb, 3, 3), tot_kap, omega, nb0, tn, tu, t_cm
    integer :: ii, j, jj, k0
    kap = reshape(kappa(1:nb * 9), [nb, 3, 3])
    write(ukap, '(1x,a)') 'la, kxx,kyy,kzz, kxy,kxz,kyz, k_RTA'
    tot_kap = 0
    do j = 1, 3
       do ii = 1, nb
          tot_kap = tot_kap + kap(ii, j, j) / 3
       end do
    end do
    do ii = 1, nb
       write(ukap, '(i2,99(1x,g11.4))') ii, (kap(ii, j, j), j = 1, 3), ((kap(ii, j, jj), jj = j + 1, 3), j = 1, 2)
    end do
    write(ukap, *) ' '
    write(ukap, *) 'tot_kap=', tot_kap
    write(ukap, *) ' '
    write(ukap, '(2(3x,2(3(1x,f9.3),2x)))') (sum(kap(:, j, j)), j = 1, 3)
    ! one header record (list-directed output may split a long string over records)
    write(utau, '(1x,a)') 'tempk,i,kibz(:,i),ii,omega,nb0,1/tauinv_N(i,ii),1/tauinv_U(i,ii),1/tauinv_tot(i,ii),' // &
         '1/tauinv_eff(i,ii),veloc(:,ii,mapinv(i))'
    t_cm = tempk * k_b / (100 * h_plank * c_light)
    do i = 1, nibz
       k0 = mapinv(i) + 1
       do ii = 1, nb
          omega = sqrt(abs(eigenval(ii, k0))) * cnst
          nb0 = nbe(omega, t_cm)
          tn = taun(i, ii)
          tu = tauu(i, ii)
          if (tu == 0) tu = 1.0d10
          if (tn == 0) tn = 1.0d10
          write(utau, '(1x,g14.8,1x,i14,3(1x,g14.8),1x,i14,99(1x,g14.8))') tempk, i, kibz(:, i), ii, omega, nb0, &
               1 / tn, 1 / tu, 1 / tautot(i, ii), taueff(i, ii), veloc(:, ii, k0)
       end do
    end do
  end subroutine write_kappa

end program alt3_kappa
