! Reader probe for the reference THERMACOND split-job files (test infrastructure, this
! container and the GPU box; built by oracle/Makefile.ref, never loaded by product code).
!
! Runs the reference's own readers on files that another program wrote in the working
! directory and dumps what they read as raw little-endian real(8):
!   read_Collision_Matrix (sub.f90:5303-5363): Cs_sycollisionm.NNN.dat under v3path -> cs.bin
!   read_CM_1D_mpi        (sub.f90:5400-5435): CM-1D.dat                            -> cm1d.bin
!   read_RHS              (sub.f90:5366-5397): RHS_IBZ.dat                          -> rhs.bin
!   read_DS               (sub.f90:5438-5469): DS.dat                               -> ds.bin
! readers.txt holds nibz, ndn, ksub_size and v3path (the module variables the readers use);
! a reader whose file is absent is skipped.
program ref_readers
 use ios
 use kpoints
 use phi3
 implicit none
 integer :: ndn, n, u
 real(8), allocatable :: cs(:,:), cm(:,:), v(:)
 logical :: ex
 character(len=99) :: fname

 open(77, file='readers.txt', status='old')
 read(77,*) nibz, ndn, ksub_size
 read(77,'(a)') v3path
 close(77)
 n = nibz*ndn*3
 write(fname,fmt="(a,i3.3,a)") "Cs_sycollisionm.", 1, ".dat"
 inquire(file=trim(v3path)//trim(fname), exist=ex)
 if (ex) then
    allocate(cs(n,n))
    cs = 0
    call read_Collision_Matrix(ndn, cs)
    call dump('cs.bin', cs, n*n)
 endif
 inquire(file='CM-1D.dat', exist=ex)
 if (ex) then
    allocate(cm(n,n))
    call read_CM_1D_mpi(nibz, ndn, cm)
    call dump('cm1d.bin', cm, n*n)
 endif
 allocate(v(n))
 inquire(file='RHS_IBZ.dat', exist=ex)
 if (ex) then
    call read_RHS(nibz, ndn, v)
    call dump('rhs.bin', v, n)
 endif
 inquire(file='DS.dat', exist=ex)
 if (ex) then
    call read_DS(nibz, ndn, v)
    call dump('ds.bin', v, n)
 endif
 write(*,'(a)') 'ref_readers: done'

contains

 subroutine dump(name, a, m)
  character(*), intent(in) :: name
  integer, intent(in) :: m
  real(8), intent(in) :: a(m)
  open(newunit=u, file=name, access='stream', form='unformatted', status='replace')
  write(u) a
  close(u)
 end subroutine dump

end program ref_readers
