!-----------------------------------------------------------------------
! refdrv.f90 -- C-callable driver over the reference ænet modules.
!
! TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile together with the
! unmodified reference sources (read in place under /root/reference/src)
! into oracle/_ref/libaenet_ref.so.  It calls the reference's own routines
! in the sequence predict.x uses them (src/predict.F90:574-706 for the
! energy loop, src/geometry.f90:661-736 for Cartesian -> lattice
! coordinates), so the numbers it returns are the reference's.
!-----------------------------------------------------------------------
module refdrv

  use, intrinsic :: iso_c_binding, only: c_int, c_double

  use constants, only: PI
  use geometry,  only: geo_recip_lattice, geo_get_bounds
  use lclist,    only: lcl_init, lcl_final, lcl_nbdist_cart
  use aenet,     only: aenet_Rc_min, aenet_Rc_max, aenet_nnb_max, &
                       aenet_atomic_energy_and_forces,            &
                       aenet_atomic_energy,                       &
                       aenet_free_atom_energy

  implicit none
  private

contains

  ! Cartesian -> lattice coordinates as geo_init_xsf does; isolated
  ! structures get the bounding box of geo_get_bounds as lattice.
  subroutine refdrv_cart_to_frac(latt, pbc, natoms, cart, frac, latt_out) &
       bind(C, name='refdrv_cart_to_frac')
    real(c_double), dimension(3,3),      intent(in)  :: latt
    integer(c_int), value,               intent(in)  :: pbc
    integer(c_int), value,               intent(in)  :: natoms
    real(c_double), dimension(3,natoms), intent(in)  :: cart
    real(c_double), dimension(3,natoms), intent(out) :: frac
    real(c_double), dimension(3,3),      intent(out) :: latt_out

    double precision, dimension(3,natoms) :: work
    double precision, dimension(3,3)      :: rec
    double precision, dimension(3)        :: origin

    work = cart
    if (pbc /= 0) then
       latt_out = latt
    else
       call geo_get_bounds(work, latt_out, origin)
    end if
    rec = geo_recip_lattice(latt_out)
    frac = matmul(rec, work)/(2.0d0*PI)

  end subroutine refdrv_cart_to_frac

  ! Neighbor lists of all atoms from lcl_init + lcl_nbdist_cart with the
  ! cutoffs of the loaded potentials.  frac is wrapped in place.
  subroutine refdrv_neighbors(latt, pbc, natoms, types, frac, cap, nnb, &
                              nbcoo, nbdist, nblist, nbtype, stat) &
       bind(C, name='refdrv_neighbors')
    real(c_double), dimension(3,3),              intent(in)    :: latt
    integer(c_int), value,                       intent(in)    :: pbc
    integer(c_int), value,                       intent(in)    :: natoms
    integer(c_int), dimension(natoms),   target, intent(in)    :: types
    real(c_double), dimension(3,natoms), target, intent(inout) :: frac
    integer(c_int), value,                       intent(in)    :: cap
    integer(c_int), dimension(natoms),           intent(out)   :: nnb
    real(c_double), dimension(3,cap,natoms),     intent(out)   :: nbcoo
    real(c_double), dimension(cap,natoms),       intent(out)   :: nbdist
    integer(c_int), dimension(cap,natoms),       intent(out)   :: nblist
    integer(c_int), dimension(cap,natoms),       intent(out)   :: nbtype
    integer(c_int),                              intent(out)   :: stat

    integer :: iatom, n, st
    logical :: pbc_f

    pbc_f = (pbc /= 0)
    stat = 0
    call lcl_init(aenet_Rc_min, aenet_Rc_max, latt, natoms, types, frac, pbc_f)
    do iatom = 1, natoms
       n = cap
       call lcl_nbdist_cart(iatom, n, nbcoo(:,:,iatom), nbdist(:,iatom), &
                            aenet_Rc_max, nblist=nblist(:,iatom), &
                            nbtype=nbtype(:,iatom), stat=st)
       if (st /= 0) then
          stat = st
          nnb(iatom) = 0
       else
          nnb(iatom) = n
       end if
    end do
    call lcl_final()

  end subroutine refdrv_neighbors

  ! Energy (and forces) of one structure given in lattice coordinates, as
  ! get_energy computes them; only atoms with mod(i-1,nranks)==rank are
  ! evaluated (the round-robin split of predict.x).
  subroutine refdrv_energy(latt, pbc, natoms, types, frac, rank, nranks, &
                           do_forces, Ecoh, Etot, Eat, F, stat) &
       bind(C, name='refdrv_energy')
    real(c_double), dimension(3,3),              intent(in)    :: latt
    integer(c_int), value,                       intent(in)    :: pbc
    integer(c_int), value,                       intent(in)    :: natoms
    integer(c_int), dimension(natoms),   target, intent(in)    :: types
    real(c_double), dimension(3,natoms), target, intent(inout) :: frac
    integer(c_int), value,                       intent(in)    :: rank
    integer(c_int), value,                       intent(in)    :: nranks
    integer(c_int), value,                       intent(in)    :: do_forces
    real(c_double),                              intent(out)   :: Ecoh
    real(c_double),                              intent(out)   :: Etot
    real(c_double), dimension(natoms),           intent(out)   :: Eat
    real(c_double), dimension(3,natoms),         intent(out)   :: F
    integer(c_int),                              intent(out)   :: stat

    integer,          dimension(aenet_nnb_max)   :: nblist
    double precision, dimension(3,aenet_nnb_max) :: nbcoo
    double precision, dimension(aenet_nnb_max)   :: nbdist
    integer,          dimension(aenet_nnb_max)   :: nbtype
    double precision, dimension(3)               :: coo_i
    double precision                             :: E_i
    integer :: iatom, nnb, type_i, st
    logical :: pbc_f

    pbc_f = (pbc /= 0)
    stat = 0
    F(1:3,1:natoms) = 0.0d0
    Eat(1:natoms) = 0.0d0
    call lcl_init(aenet_Rc_min, aenet_Rc_max, latt, natoms, types, frac, pbc_f)
    Ecoh = 0.0d0
    Etot = 0.0d0
    do iatom = 1, natoms
       if (mod(iatom-1, nranks) /= rank) cycle
       type_i = types(iatom)
       coo_i(1:3) = matmul(latt, frac(1:3,iatom))
       nnb = aenet_nnb_max
       call lcl_nbdist_cart(iatom, nnb, nbcoo, nbdist, aenet_Rc_max, &
                            nblist=nblist, nbtype=nbtype)
       if (do_forces /= 0) then
          call aenet_atomic_energy_and_forces(coo_i, type_i, iatom, nnb, &
               nbcoo, nbtype, nblist, natoms, E_i, F, st)
       else
          call aenet_atomic_energy(coo_i, type_i, nnb, nbcoo, nbtype, E_i, st)
       end if
       if (st /= 0) stat = st
       Etot = Etot + E_i
       Ecoh = Ecoh + E_i - aenet_free_atom_energy(type_i)
       Eat(iatom) = E_i
    end do
    call lcl_final()

  end subroutine refdrv_energy

  ! Timed evaluation of the atoms with mod(i-1,nranks)==rank, for the CPU
  ! baseline: t_nbl = lcl_init plus the neighbor lists of those atoms
  ! (lcl_nbdist_cart), t_eval = aenet_atomic_energy_and_forces on the
  ! stored lists (the "kernel-equivalent" part); wall-clock seconds.
  subroutine refdrv_timed(latt, pbc, natoms, types, frac, rank, nranks, &
                          t_nbl, t_eval, Etot, E_out, F_out, stat) bind(C, name='refdrv_timed')
    real(c_double), dimension(3,3),              intent(in)    :: latt
    integer(c_int), value,                       intent(in)    :: pbc
    integer(c_int), value,                       intent(in)    :: natoms
    integer(c_int), dimension(natoms),   target, intent(in)    :: types
    real(c_double), dimension(3,natoms), target, intent(inout) :: frac
    integer(c_int), value,                       intent(in)    :: rank
    integer(c_int), value,                       intent(in)    :: nranks
    real(c_double),                              intent(out)   :: t_nbl
    real(c_double),                              intent(out)   :: t_eval
    real(c_double),                              intent(out)   :: Etot
    ! energies of this rank's atoms (others untouched) and the force sums of
    ! their contributions to every atom (the reference's F of these calls)
    real(c_double), dimension(natoms),           intent(inout) :: E_out
    real(c_double), dimension(3,natoms),         intent(out)   :: F_out
    integer(c_int),                              intent(out)   :: stat

    integer,          dimension(:,:),   allocatable :: nblist, nbtype
    double precision, dimension(:,:,:), allocatable :: nbcoo
    double precision, dimension(:,:),   allocatable :: nbdist, cooi
    integer,          dimension(:),     allocatable :: nnb, idx
    double precision, dimension(:,:),   allocatable :: F
    double precision :: E_i
    integer :: iatom, k, nsub, st
    integer(8) :: c0, c1, crate
    logical :: pbc_f

    pbc_f = (pbc /= 0)
    stat = 0
    nsub = 0
    do iatom = 1, natoms
       if (mod(iatom-1, nranks) == rank) nsub = nsub + 1
    end do
    allocate(nblist(aenet_nnb_max,nsub), nbtype(aenet_nnb_max,nsub), &
             nbcoo(3,aenet_nnb_max,nsub), nbdist(aenet_nnb_max,nsub), &
             cooi(3,nsub), nnb(nsub), idx(nsub), F(3,natoms))
    call system_clock(c0, crate)
    call lcl_init(aenet_Rc_min, aenet_Rc_max, latt, natoms, types, frac, pbc_f)
    k = 0
    do iatom = 1, natoms
       if (mod(iatom-1, nranks) /= rank) cycle
       k = k + 1
       idx(k) = iatom
       cooi(1:3,k) = matmul(latt, frac(1:3,iatom))
       nnb(k) = aenet_nnb_max
       call lcl_nbdist_cart(iatom, nnb(k), nbcoo(:,:,k), nbdist(:,k), aenet_Rc_max, &
                            nblist=nblist(:,k), nbtype=nbtype(:,k))
    end do
    call lcl_final()
    call system_clock(c1)
    t_nbl = dble(c1 - c0)/dble(crate)
    F = 0.0d0
    Etot = 0.0d0
    call system_clock(c0)
    do k = 1, nsub
       call aenet_atomic_energy_and_forces(cooi(:,k), types(idx(k)), idx(k), nnb(k), &
            nbcoo(:,1:nnb(k),k), nbtype(1:nnb(k),k), nblist(1:nnb(k),k), natoms, E_i, F, st)
       if (st /= 0) stat = st
       Etot = Etot + E_i
       E_out(idx(k)) = E_i
    end do
    call system_clock(c1)
    t_eval = dble(c1 - c0)/dble(crate)
    F_out = F
    deallocate(nblist, nbtype, nbcoo, nbdist, cooi, nnb, idx, F)

  end subroutine refdrv_timed

end module refdrv
