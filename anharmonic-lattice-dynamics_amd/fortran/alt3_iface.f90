!  iso_c_binding interface to the alt3 C ABI (include/alt3.h).
!
!  Each interface binds one extern "C" entry point.  Array arguments are passed as
!  type(c_ptr) so that a caller can hand over c_loc(array) or c_null_ptr ("keep on the
!  device"); the arrays keep THERMACOND's column-major shapes, e.g. eigenval(ndyn,nkc),
!  Collision_Matrixmpi(ksub*ndyn*3, nibz*ndyn*3), kappa(ndyn,3,3).
!  The subroutines these calls replace are cited in include/alt3.h.
module alt3_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: ALT3_OK = 0
  integer(c_int), parameter :: ALT3_V3_FP64_REF = 0, ALT3_V3_FP64 = 1, ALT3_V3_MIXED = 2

  ! alt3_system of include/alt3.h, field for field
  type, bind(C) :: alt3_system
     integer(c_int) :: natoms0
     type(c_ptr) :: mass, r123, latparams, primlatt
     integer(c_int) :: natoms
     type(c_ptr) :: atompos, iatomcell0
     integer(c_int) :: nt2
     type(c_ptr) :: fc2_atoms, fc2_xyz, fc2
     integer(c_int) :: nt3
     type(c_ptr) :: fc3_atoms, fc3_xyz, fc3
     integer(c_int) :: nc(3)
     integer(c_int) :: classical, host_only
     integer(c_int) :: born_flag
     type(c_ptr) :: epsil, zeu
  end type alt3_system

  interface
     integer(c_int) function alt3_create(sys, ctx) bind(C, name='alt3_create')
       import :: c_int, c_ptr, alt3_system
       type(alt3_system), intent(in) :: sys
       type(c_ptr), intent(out) :: ctx
     end function alt3_create

     integer(c_int) function alt3_destroy(ctx) bind(C, name='alt3_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function alt3_destroy

     type(c_ptr) function alt3_last_error() bind(C, name='alt3_last_error')
       import :: c_ptr
     end function alt3_last_error

     integer(c_int) function alt3_sizes(ctx, nk, nibz, nb, nops) bind(C, name='alt3_sizes')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), intent(out) :: nk, nibz, nb, nops
     end function alt3_sizes

     integer(c_int) function alt3_get_mesh(ctx, kpc_eig, kpc, mapibz, mapinv, kibz, wibz, tet, opk, mappos, npos) &
          bind(C, name='alt3_get_mesh')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, kpc_eig, kpc, mapibz, mapinv, kibz, wibz, tet, opk, mappos, npos
     end function alt3_get_mesh

     integer(c_int) function alt3_eig(ctx, eigenval, eigenvec, veloc) bind(C, name='alt3_eig')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, eigenval, eigenvec, veloc
     end function alt3_eig
     ! load eigen data computed elsewhere (same layouts as alt3_eig)
     integer(c_int) function alt3_set_eig(ctx, eigenval, eigenvec, veloc) bind(C, name='alt3_set_eig')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, eigenval, eigenvec, veloc
     end function alt3_set_eig

     integer(c_int) function alt3_v3sq_ibz(ctx, ibz_lo, ibz_hi, precision, v33s8) bind(C, name='alt3_v3sq_ibz')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, v33s8
       integer(c_int), value :: ibz_lo, ibz_hi, precision
     end function alt3_v3sq_ibz

     integer(c_int) function alt3_fgr(ctx, p1sm, p2sm) bind(C, name='alt3_fgr')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, p1sm, p2sm
     end function alt3_fgr

     integer(c_int) function alt3_tau(ctx, tempk, q, qn, qu, taun, tauu, tautot) bind(C, name='alt3_tau')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, q, qn, qu, taun, tauu, tautot
       real(c_double), value :: tempk
     end function alt3_tau

     integer(c_int) function alt3_kappa_rta(ctx, tempk, kappa, frta, taueff) bind(C, name='alt3_kappa_rta')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, kappa, frta, taueff
       real(c_double), value :: tempk
     end function alt3_kappa_rta

     integer(c_int) function alt3_collision(ctx, tempk, cm, rhs) bind(C, name='alt3_collision')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, cm, rhs
       real(c_double), value :: tempk
     end function alt3_collision

     integer(c_int) function alt3_cg(ctx, n, a, b, x, max_iter, ndig, funcmin, iters) bind(C, name='alt3_cg')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, a, b, x, funcmin, iters
       integer(c_int), value :: n, max_iter, ndig
     end function alt3_cg

     integer(c_int) function alt3_kappa_bte(ctx, tempk, x, kappa, taueff) bind(C, name='alt3_kappa_bte')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, x, kappa, taueff
       real(c_double), value :: tempk
     end function alt3_kappa_bte

     integer(c_int) function alt3_kappa_iter(ctx, tempk, max_iter, update_mix, tol, kappa, kappa_hist, taueff_hist, &
          fff, iters) bind(C, name='alt3_kappa_iter')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, kappa, kappa_hist, taueff_hist, fff, iters
       real(c_double), value :: tempk, update_mix, tol
       integer(c_int), value :: max_iter
     end function alt3_kappa_iter

     integer(c_int) function alt3_self_w3(ctx, nq, q, tempk, eta, v3_threshold, omega0, nself, uself) &
          bind(C, name='alt3_self_w3')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, q, omega0, nself, uself
       integer(c_int), value :: nq
       real(c_double), value :: tempk, eta, v3_threshold
     end function alt3_self_w3
     integer(c_int) function alt3_self_w4(ctx, nq, q, tempk, eta, v3_threshold, omega0, nself, uself) &
          bind(C, name='alt3_self_w4')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, q, omega0, nself, uself
       integer(c_int), value :: nq
       real(c_double), value :: tempk, eta, v3_threshold
     end function alt3_self_w4

     integer(c_int) function alt3_update(ctx, tempk, fff, frta, f2) bind(C, name='alt3_update')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, fff, frta, f2
       real(c_double), value :: tempk
     end function alt3_update

     integer(c_int) function alt3_set_iso(ctx, giso, piso) bind(C, name='alt3_set_iso')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, giso, piso
     end function alt3_set_iso

     integer(c_int) function alt3_isotope_g(element, g) bind(C, name='alt3_isotope_g')
       import :: c_int, c_char, c_double
       character(kind=c_char), dimension(*) :: element
       real(c_double) :: g
     end function alt3_isotope_g

     integer(c_int) function alt3_solve(ctx, ntemps, temps, precision, kappa_rta, kappa_bte, cg_iters, stage_ms) &
          bind(C, name='alt3_solve')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, temps, kappa_rta, kappa_bte, cg_iters, stage_ms
       integer(c_int), value :: ntemps, precision
     end function alt3_solve
  end interface

contains

  ! The reference writes a log line and stops on an error; so does this check.
  subroutine alt3_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (rc == ALT3_OK) return
    p = alt3_last_error()
    n = 0
    if (c_associated(p)) then
       call c_f_pointer(p, msg, [4096])
       do while (n < 4096)
          if (msg(n + 1) == c_null_char) exit
          n = n + 1
       end do
       write(*, '(a,a,i0,a,4096a)') where, ': alt3 error ', rc, ': ', msg(1:n)
    else
       write(*, '(a,a,i0)') where, ': alt3 error ', rc
    end if
    stop 1
  end subroutine alt3_check

end module alt3_iface
