!
! Checker-only harness: drives the reference's own Math_Layer modules
! (Legendre_Polynomials, Legendre_Transforms, Chebyshev_Polynomials,
! Spectral_Derivatives, Structures),
! compiled from their sources under the reference tree, through bind(C) entry
! points that Python (ctypes) calls when it validates the C restatement in
! oracle/raysht_oracle.c and writes the golden fixtures under tests/golden/.
!
! This is test infrastructure.  Nothing in the product links or calls it.
!
! Layouts (all column-major, as the reference stores them):
!   spectral, packed over mp = 1..n_m in m_values order:
!       s(mp)%data(m:l_max, 1:nr, 1:2, 1:nf)      l fastest
!   hybrid p2(1:n_theta, 1:2*nr*nf, 1:n_m)        theta fastest
!   Chebyshev buffers f(1:n, 1:ncols)             radius fastest
!
Module raysht_ref_harness
    Use ISO_C_BINDING
    Use Structures
    Use Legendre_Polynomials
    Use Legendre_Transforms
    Use Chebyshev_Polynomials, Only : Cheby_Grid
    Use Spectral_Derivatives
    Implicit None
    Type(Cheby_Grid), Save, Target :: h_grid
    Logical, Save :: h_grid_ready = .false.
    Integer, Save :: h_ncheb = 0
    Integer, Save :: h_ang_nm = 0, h_ang_lmax = 0, h_ang_nr = 0
    Integer, Allocatable, Save :: h_ang_mvals(:)

Contains

    Subroutine release_legendre()
        Integer :: i
        If (.not. allocated(coloc)) Return
        ! Finalize_Legendre (Legendre_Polynomials.F90:58-63) leaves the
        ! integration-weighted parity tables allocated; release them here so
        ! that Initialize_Legendre can run again in the same process.
        If (allocated(ip_lm_even)) Then
            Do i = 1, size(ip_lm_even)
                If (allocated(ip_lm_even(i)%data)) DeAllocate(ip_lm_even(i)%data)
            Enddo
            DeAllocate(ip_lm_even)
        Endif
        If (allocated(ip_lm_odd)) Then
            Do i = 1, size(ip_lm_odd)
                If (allocated(ip_lm_odd(i)%data)) DeAllocate(ip_lm_odd(i)%data)
            Enddo
            DeAllocate(ip_lm_odd)
        Endif
        Call Finalize_Legendre()
    End Subroutine release_legendre

    Subroutine ref_legendre_init(nt, lmax, nm, mvals) bind(C, name='ref_legendre_init')
        Integer(C_INT), Value :: nt, lmax, nm
        Integer(C_INT), Intent(In) :: mvals(nm)
        Integer :: mv(nm)
        Call release_legendre()
        mv(:) = mvals(:)
        Call Initialize_Legendre(nt, lmax, mv, .true.)
    End Subroutine ref_legendre_init

    ! Gauss-Legendre nodes and weights (Real*16) exported as head/tail double pairs.
    Subroutine ref_legendre_nodes(xhi, xlo, whi, wlo) bind(C, name='ref_legendre_nodes')
        Real(C_DOUBLE), Intent(Out) :: xhi(n_theta), xlo(n_theta), whi(n_theta), wlo(n_theta)
        Integer :: i
        Real*16 :: q
        Do i = 1, n_theta
            xhi(i) = Real(coloc(i), 8)
            q = coloc(i) - Real(xhi(i), 16)
            xlo(i) = Real(q, 8)
            whi(i) = Real(gl_weights(i), 8)
            q = gl_weights(i) - Real(whi(i), 16)
            wlo(i) = Real(q, 8)
        Enddo
    End Subroutine ref_legendre_nodes

    ! The same nodes and weights as stored (Real*16), 16 bytes per value.
    Subroutine ref_legendre_nodes_q(px, pw) bind(C, name='ref_legendre_nodes_q')
        Type(C_PTR), Value :: px, pw
        Real*16, Pointer :: x16(:), w16(:)
        Call c_f_pointer(px, x16, [n_theta])
        Call c_f_pointer(pw, w16, [n_theta])
        x16(:) = coloc(1:n_theta)
        w16(:) = gl_weights(1:n_theta)
    End Subroutine ref_legendre_nodes_q

    ! The four parity tables of every mp, concatenated in mp order, each in its
    ! stored orientation: p_lm_even(mp)%data(1:n_e,1:nt2) (l fastest) and
    ! ip_lm_even(mp)%data(1:nt2,1:n_e) (theta fastest); odd twins likewise.
    Subroutine ref_legendre_tables(pe, po, ipe, ipo) bind(C, name='ref_legendre_tables')
        Real(C_DOUBLE), Intent(Out) :: pe(*), po(*), ipe(*), ipo(*)
        Integer :: mp, ne, no, nt2, oe, oo, k
        nt2 = n_theta/2
        oe = 0
        oo = 0
        Do mp = 1, n_m
            ne = n_l_even(mp)
            no = n_l_odd(mp)
            k = ne*nt2
            If (ne > 0) Then
                pe(oe+1:oe+k) = Reshape(p_lm_even(mp)%data, (/ k /))
                ipe(oe+1:oe+k) = Reshape(ip_lm_even(mp)%data, (/ k /))
            Endif
            oe = oe + k
            k = no*nt2
            If (no > 0) Then
                po(oo+1:oo+k) = Reshape(p_lm_odd(mp)%data, (/ k /))
                ipo(oo+1:oo+k) = Reshape(ip_lm_odd(mp)%data, (/ k /))
            Endif
            oo = oo + k
        Enddo
    End Subroutine ref_legendre_tables

    ! StP through the reference's generic Legendre_Transform (resolves to StP_4d_dgp2).
    Subroutine ref_stp(nr, nf, s_in, p_out) bind(C, name='ref_stp')
        Integer(C_INT), Value :: nr, nf
        Real(C_DOUBLE), Intent(In) :: s_in(*)
        Real(C_DOUBLE), Intent(Out) :: p_out(n_theta, 2*nr*nf, n_m)
        Type(rmcontainer4d), Allocatable :: s(:)
        Integer :: mp, m, n, off
        Allocate(s(1:n_m))
        off = 0
        Do mp = 1, n_m
            m = m_values(mp)
            Allocate(s(mp)%data(m:l_max, 1:nr, 1:2, 1:nf))
            n = (l_max-m+1)*nr*2*nf
            s(mp)%data = Reshape(s_in(off+1:off+n), (/ l_max-m+1, nr, 2, nf /))
            off = off + n
        Enddo
        Call Legendre_Transform(s, p_out)
        Do mp = 1, n_m
            DeAllocate(s(mp)%data)
        Enddo
        DeAllocate(s)
    End Subroutine ref_stp

    ! PtS through the generic Legendre_Transform (resolves to PtS_4d_dgpv3).
    Subroutine ref_pts(nr, nf, p_in, s_out) bind(C, name='ref_pts')
        Integer(C_INT), Value :: nr, nf
        Real(C_DOUBLE), Intent(In) :: p_in(n_theta, 2*nr*nf, n_m)
        Real(C_DOUBLE), Intent(Out) :: s_out(*)
        Type(rmcontainer4d), Allocatable :: s(:)
        Integer :: mp, m, n, off
        Allocate(s(1:n_m))
        Do mp = 1, n_m
            m = m_values(mp)
            Allocate(s(mp)%data(m:l_max, 1:nr, 1:2, 1:nf))
            s(mp)%data = 0.0d0
        Enddo
        Call Legendre_Transform(p_in, s)
        off = 0
        Do mp = 1, n_m
            m = m_values(mp)
            n = (l_max-m+1)*nr*2*nf
            s_out(off+1:off+n) = Reshape(s(mp)%data, (/ n /))
            off = off + n
            DeAllocate(s(mp)%data)
        Enddo
        DeAllocate(s)
    End Subroutine ref_pts

    ! One-domain Chebyshev grid on [rmin, rmax] with n points (zeros grid).  dealias_by is always
    ! passed, 0 (the default rda) here: Initialize_Cheby_Grid's custom_dealiasing is a local with
    ! an initializer, hence saved between calls, so a call without the argument after one with
    ! it would read the absent argument.
    Subroutine ref_cheby_init(n, rmin, rmax, grid, weights) bind(C, name='ref_cheby_init')
        Integer(C_INT), Value :: n
        Real(C_DOUBLE), Value :: rmin, rmax
        Real(C_DOUBLE), Intent(Out) :: grid(n), weights(n)
        Integer :: npoly(1), db(1)
        Real*8 :: bounds(2)
        Type(Cheby_Grid) :: fresh
        npoly(1) = n
        bounds(1) = rmin
        bounds(2) = rmax
        db(1) = 0
        h_grid = fresh
        Call h_grid%init(grid, weights, 1, npoly, bounds, dealias_by=db)
        h_grid_ready = .true.
        h_ncheb = n
    End Subroutine ref_cheby_init

    ! Cheby_Grid%init with nd domains: npoly(nd), bounds(nd + 1) ascending, dealias_by(nd) when
    ! has_db /= 0, else zeros (the default rda; passed always, as in ref_cheby_init); ntotal =
    ! sum(npoly) radii into grid and weights.
    Subroutine ref_cheby_init_domains(nd, npoly, bounds, has_db, dealias_by, ntotal, grid, weights) &
            bind(C, name='ref_cheby_init_domains')
        Integer(C_INT), Value :: nd, has_db, ntotal
        Integer(C_INT), Intent(In) :: npoly(nd), dealias_by(nd)
        Real(C_DOUBLE), Intent(In) :: bounds(nd + 1)
        Real(C_DOUBLE), Intent(Out) :: grid(ntotal), weights(ntotal)
        Integer :: np(nd), db(nd)
        Real*8 :: bd(nd + 1)
        Type(Cheby_Grid) :: fresh
        np = npoly
        db = dealias_by
        bd = bounds
        If (has_db == 0) db = 0
        h_grid = fresh
        Call h_grid%init(grid, weights, nd, np, bd, dealias_by=db)
        h_grid_ready = .true.
        h_ncheb = ntotal
    End Subroutine ref_cheby_init_domains

    ! the grid's deriv_scaling(1:ntotal) and rda(1:domain_count)
    Subroutine ref_cheby_grid_arrays(dscale, rda) bind(C, name='ref_cheby_grid_arrays')
        Real(C_DOUBLE), Intent(Out) :: dscale(*)
        Integer(C_INT), Intent(Out) :: rda(*)
        dscale(1:h_grid%ntotal) = h_grid%deriv_scaling(1:h_grid%ntotal)
        rda(1:h_grid%domain_count) = h_grid%rda(1:h_grid%domain_count)
    End Subroutine ref_cheby_grid_arrays

    ! dealias_buffer4d on a (n, ncols, 1, 1) buffer, in place
    Subroutine ref_cheby_dealias(ncols, buf) bind(C, name='ref_cheby_dealias')
        Integer(C_INT), Value :: ncols
        Real(C_DOUBLE), Intent(InOut) :: buf(h_ncheb, ncols, 1, 1)
        Call h_grid%dealias_buffer(buf)
    End Subroutine ref_cheby_dealias

    Subroutine ref_cheby_tables(ceven, codd) bind(C, name='ref_cheby_tables')
        Real(C_DOUBLE), Intent(Out) :: ceven(*), codd(*)
        Integer :: ke, ko
        ke = size(h_grid%cheby_even(1)%data)
        ko = size(h_grid%cheby_odd(1)%data)
        ceven(1:ke) = Reshape(h_grid%cheby_even(1)%data, (/ ke /))
        codd(1:ko) = Reshape(h_grid%cheby_odd(1)%data, (/ ko /))
    End Subroutine ref_cheby_tables

    ! To_Spectral4D on a (n, ncols, 1, 1) buffer.
    Subroutine ref_cheby_to_spectral(ncols, f_in, c_out) bind(C, name='ref_cheby_to_spectral')
        Integer(C_INT), Value :: ncols
        Real(C_DOUBLE), Intent(In) :: f_in(h_ncheb, ncols, 1, 1)
        Real(C_DOUBLE), Intent(Out) :: c_out(h_ncheb, ncols, 1, 1)
        Call h_grid%to_spectral(f_in, c_out)
    End Subroutine ref_cheby_to_spectral

    ! From_Spectral4D on a (n, ncols, 1, 1) buffer.
    Subroutine ref_cheby_from_spectral(ncols, c_in, f_out) bind(C, name='ref_cheby_from_spectral')
        Integer(C_INT), Value :: ncols
        Real(C_DOUBLE), Intent(In) :: c_in(h_ncheb, ncols, 1, 1)
        Real(C_DOUBLE), Intent(Out) :: f_out(h_ncheb, ncols, 1, 1)
        Call h_grid%from_spectral(c_in, f_out)
    End Subroutine ref_cheby_from_spectral

    ! Cheby_Deriv_Buffer_4D (d_by_dr_cp) out of place: buffer(:,:,1,1) = c_in,
    ! buffer(:,:,1,2) receives the dorder-th derivative's coefficients.
    Subroutine ref_cheby_deriv(ncols, dorder, c_in, c_out) bind(C, name='ref_cheby_deriv')
        Integer(C_INT), Value :: ncols, dorder
        Real(C_DOUBLE), Intent(In) :: c_in(h_ncheb, ncols)
        Real(C_DOUBLE), Intent(Out) :: c_out(h_ncheb, ncols)
        Real*8, Allocatable :: buffer(:,:,:,:)
        Allocate(buffer(h_ncheb, ncols, 1, 2))
        buffer(:,:,1,1) = c_in
        buffer(:,:,1,2) = 0.0d0
        Call h_grid%d_by_dr_cp(1, 2, buffer, dorder)
        c_out = buffer(:,:,1,2)
        DeAllocate(buffer)
    End Subroutine ref_cheby_deriv

    ! Cheby_Deriv_OutOf_Buffer (d_by_dr_nobuffer): the same recursion into a separate array.
    Subroutine ref_cheby_deriv_nobuffer(ncols, dorder, c_in, c_out) bind(C, name='ref_cheby_deriv_nobuffer')
        Integer(C_INT), Value :: ncols, dorder
        Real(C_DOUBLE), Intent(In) :: c_in(h_ncheb, ncols)
        Real(C_DOUBLE), Intent(Out) :: c_out(h_ncheb, ncols)
        Real*8, Allocatable :: buffer(:,:,:,:), obuffer(:,:,:)
        Allocate(buffer(h_ncheb, ncols, 1, 1), obuffer(h_ncheb, ncols, 1))
        buffer(:,:,1,1) = c_in
        Call h_grid%d_by_dr_nobuffer(1, buffer, obuffer, dorder)
        c_out = obuffer(:,:,1)
        DeAllocate(buffer, obuffer)
    End Subroutine ref_cheby_deriv_nobuffer

    ! Initialize_Angular_Derivatives (Spectral_Derivatives.F90:49-83) for nm m values,
    ! l_max and nr radii (a fresh start each call).
    Subroutine ref_angular_init(nm, mvals, lmax, nr) bind(C, name='ref_angular_init')
        Integer(C_INT), Value :: nm, lmax, nr
        Integer(C_INT), Intent(In) :: mvals(nm)
        If (Allocated(mlocal)) DeAllocate(mlocal)
        If (Allocated(deriv_coefs)) DeAllocate(deriv_coefs)
        If (Allocated(sderiv_coefs)) DeAllocate(sderiv_coefs)
        If (Allocated(h_ang_mvals)) DeAllocate(h_ang_mvals)
        Allocate(h_ang_mvals(nm))
        h_ang_mvals = mvals
        h_ang_nm = nm
        h_ang_lmax = lmax
        h_ang_nr = nr
        Call Initialize_Angular_Derivatives(h_ang_mvals, lmax, nr)
    End Subroutine ref_angular_init

    ! packed spectral blocks of nf fields <-> rmcontainer4d / rmcontainer3d
    Subroutine unpack4d(buf, nf, a)
        Real(C_DOUBLE), Intent(In) :: buf(*)
        Integer, Intent(In) :: nf
        Type(rmcontainer4d), Allocatable, Intent(Out) :: a(:)
        Integer :: mp, m, n, off
        Allocate(a(1:h_ang_nm))
        off = 0
        Do mp = 1, h_ang_nm
            m = h_ang_mvals(mp)
            Allocate(a(mp)%data(m:h_ang_lmax, 1:h_ang_nr, 1:2, 1:nf))
            n = (h_ang_lmax-m+1)*h_ang_nr*2*nf
            a(mp)%data = Reshape(buf(off+1:off+n), (/ h_ang_lmax-m+1, h_ang_nr, 2, nf /))
            off = off + n
        Enddo
    End Subroutine unpack4d

    Subroutine pack4d(a, nf, buf)
        Type(rmcontainer4d), Intent(In) :: a(:)
        Integer, Intent(In) :: nf
        Real(C_DOUBLE), Intent(InOut) :: buf(*)
        Integer :: mp, m, n, off
        off = 0
        Do mp = 1, h_ang_nm
            m = h_ang_mvals(mp)
            n = (h_ang_lmax-m+1)*h_ang_nr*2*nf
            buf(off+1:off+n) = Reshape(a(mp)%data, (/ n /))
            off = off + n
        Enddo
    End Subroutine pack4d

    Subroutine alloc3d(a)
        Type(rmcontainer3d), Allocatable, Intent(Out) :: a(:)
        Integer :: mp, m
        Allocate(a(1:h_ang_nm))
        Do mp = 1, h_ang_nm
            m = h_ang_mvals(mp)
            Allocate(a(mp)%data(m:h_ang_lmax, 1:h_ang_nr, 1:2))
            a(mp)%data = 0.0d0
        Enddo
    End Subroutine alloc3d

    Subroutine pack3d(a, buf)
        Type(rmcontainer3d), Intent(In) :: a(:)
        Real(C_DOUBLE), Intent(Out) :: buf(*)
        Integer :: mp, m, n, off
        off = 0
        Do mp = 1, h_ang_nm
            m = h_ang_mvals(mp)
            n = (h_ang_lmax-m+1)*h_ang_nr*2
            buf(off+1:off+n) = Reshape(a(mp)%data, (/ n /))
            off = off + n
        Enddo
    End Subroutine pack3d

    ! which = 0: d_by_dtheta, 1: d_by_sdtheta, 2: d_by_dphi on a packed buffer of nf fields
    ! (in place: buf is read and written).  fout >= 1: the buffer forms (d_dtheta_buffer4D,
    ! d_sdtheta_buffer4d, d_by_dphi_rlmbuff4d: field fin -> field fout, 1-based);
    ! fout = 0: the 4dbuff2arr forms into arr, a packed one-field array.
    Subroutine ref_angular_deriv(which, nf, fin, fout, buf, arr) bind(C, name='ref_angular_deriv')
        Integer(C_INT), Value :: which, nf, fin, fout
        Real(C_DOUBLE), Intent(InOut) :: buf(*)
        Real(C_DOUBLE), Intent(Out) :: arr(*)
        Type(rmcontainer4d), Allocatable :: a(:)
        Type(rmcontainer3d), Allocatable :: b(:)
        Call unpack4d(buf, nf, a)
        If (fout >= 1) Then
            If (which == 0) Call d_by_dtheta(a, fin, fout)
            If (which == 1) Call d_by_sdtheta(a, fin, fout)
            If (which == 2) Call d_by_dphi(a, fin, fout)
            Call pack4d(a, nf, buf)
        Else
            Call alloc3d(b)
            If (which == 0) Call d_by_dtheta(a, fin, b)
            If (which == 1) Call d_by_sdtheta(a, fin, b)
            If (which == 2) Call d_by_dphi(a, fin, b)
            Call pack3d(b, arr)
            DeAllocate(b)
        Endif
        DeAllocate(a)
    End Subroutine ref_angular_deriv

    ! rlm_spaceb's hydro steps (Physics/Sphere_Hybrid_Space.F90:127-214; that module needs
    ! the whole physics state, so the harness issues the same calls of the reference's
    ! d_by_sdtheta / d_by_dphi and the same array statements itself) on a packed buffer of nf
    ! fields, in place; wv, pv, zv are 1-based.
    Subroutine ref_rlm_spaceb(nf, wv, pv, zv, buf) bind(C, name='ref_rlm_spaceb')
        Integer(C_INT), Value :: nf, wv, pv, zv
        Real(C_DOUBLE), Intent(InOut) :: buf(*)
        Type(rmcontainer4d), Allocatable :: a(:)
        Type(rmcontainer3d), Allocatable :: t1(:), t2(:)
        Real*8, Allocatable :: ol(:)
        Real*8 :: ell
        Integer :: mp, m, l, r, imi, lm
        lm = h_ang_lmax
        Allocate(ol(0:lm))
        ol(0) = 1.0d0
        Do l = 1, lm
            ell = l*1.0d0
            ol(l) = 1.0d0/(ell*(ell+1))
        Enddo
        Call unpack4d(buf, nf, a)
        Call alloc3d(t1)
        Call alloc3d(t2)
        Do mp = 1, h_ang_nm
            m = h_ang_mvals(mp)
            Do imi = 1, 2
                Do r = 1, h_ang_nr
                    a(mp)%data(m:lm, r, imi, wv) = a(mp)%data(m:lm, r, imi, wv)*ol(m:lm)
                Enddo
            Enddo
        Enddo
        Call d_by_sdtheta(a, zv, t1)
        Call d_by_dphi(a, pv, t2)
        Do mp = 1, h_ang_nm
            m = h_ang_mvals(mp)
            Do imi = 1, 2
                Do r = 1, h_ang_nr
                    t1(mp)%data(m:lm, r, imi) = (t2(mp)%data(m:lm, r, imi) - t1(mp)%data(m:lm, r, imi))*ol(m:lm)
                Enddo
            Enddo
        Enddo
        Call d_by_dphi(a, zv, t2)
        Call d_by_sdtheta(a, pv, zv)
        Do mp = 1, h_ang_nm
            m = h_ang_mvals(mp)
            Do imi = 1, 2
                Do r = 1, h_ang_nr
                    a(mp)%data(m:lm, r, imi, pv) = (a(mp)%data(m:lm, r, imi, zv) + t2(mp)%data(m:lm, r, imi))*ol(m:lm)
                    a(mp)%data(m:lm, r, imi, zv) = t1(mp)%data(m:lm, r, imi)
                Enddo
            Enddo
            If (m == 0) Then
                a(mp)%data(0, :, :, pv) = 0.0d0
                a(mp)%data(0, :, :, wv) = 0.0d0
                a(mp)%data(0, :, :, zv) = 0.0d0
            Endif
            a(mp)%data(lm, :, :, :) = 0.0d0
        Enddo
        Call pack4d(a, nf, buf)
        DeAllocate(a, t1, t2, ol)
    End Subroutine ref_rlm_spaceb

    ! rlm_spacea's Velocity_Components and Velocity_Derivatives, then d_by_dtheta(tvar, dtdt)
    ! (Physics/Sphere_Hybrid_Space.F90:80,235-355; the module needs the whole physics state,
    ! so the harness issues the same calls of the reference's d_by_dtheta / d_by_dphi and the
    ! same array statements itself) on a packed buffer of nf fields, in place.  v = 1-based
    ! (W, dWdr, d2Wdr2, Z, dZdr, T, vtheta, vphi, dvtdr, dvpdr); vr, dvrdr, dtdt and dvrdt are
    ! the W, dWdr, d2Wdr2 and dZdr fields (Fields.F90:323-326).  pseudo /= 0 adds the
    ! dsdr_over_cp term.
    Subroutine ref_rlm_spacea_hydro(nf, v, pseudo, over_rhor, over_rhorsq, drho_term, two_over_r, dlnrho, &
            dsdr_over_cp, buf) bind(C, name='ref_rlm_spacea_hydro')
        Integer(C_INT), Value :: nf, pseudo
        Integer(C_INT), Intent(In) :: v(10)
        Real(C_DOUBLE), Intent(In) :: over_rhor(*), over_rhorsq(*), drho_term(*), two_over_r(*), dlnrho(*)
        Real(C_DOUBLE), Intent(In) :: dsdr_over_cp(*)
        Real(C_DOUBLE), Intent(InOut) :: buf(*)
        Type(rmcontainer4d), Allocatable :: a(:)
        Type(rmcontainer3d), Allocatable :: ftemp1(:), ftemp2(:)
        Real*8, Allocatable :: l_l_plus1(:)
        Real*8 :: ell
        Integer :: mp, m, l, r, imi, lm
        Integer :: vr, dwdr, d2wdr2, zvar, dzdr, tvar, vtheta, vphi, dvtdr, dvpdr, dvrdr, dvrdt, dtdt
        vr = v(1); dwdr = v(2); d2wdr2 = v(3); zvar = v(4); dzdr = v(5); tvar = v(6)
        vtheta = v(7); vphi = v(8); dvtdr = v(9); dvpdr = v(10)
        dvrdr = dwdr; dtdt = d2wdr2; dvrdt = dzdr
        lm = h_ang_lmax
        Allocate(l_l_plus1(0:lm))
        Do l = 0, lm
            ell = l*1.0d0
            l_l_plus1(l) = ell*(ell+1)
        Enddo
        Call unpack4d(buf, nf, a)
        Call alloc3d(ftemp1)
        Call alloc3d(ftemp2)
        ! Velocity_Components
        Do mp = 1, h_ang_nm
            m = h_ang_mvals(mp)
            Do imi = 1, 2
                Do r = 1, h_ang_nr
                    a(mp)%data(m:lm,r,imi,vr) = l_l_plus1(m:lm)*a(mp)%data(m:lm,r,imi,vr)*over_rhorsq(r)
                Enddo
            Enddo
        Enddo
        Call d_by_dtheta(a, dwdr, ftemp1)
        Call d_by_dphi(a, zvar, ftemp2)
        Do mp = 1, h_ang_nm
            ftemp1(mp)%data = ftemp1(mp)%data+ftemp2(mp)%data
            Do r = 1, h_ang_nr
                a(mp)%data(:,r,:,vtheta) = ftemp1(mp)%data(:,r,:)*over_rhor(r)
            Enddo
        Enddo
        Call d_by_dphi(a, dwdr, ftemp1)
        Call d_by_dtheta(a, zvar, ftemp2)
        Do mp = 1, h_ang_nm
            ftemp1(mp)%data = ftemp1(mp)%data-ftemp2(mp)%data
            Do r = 1, h_ang_nr
                a(mp)%data(:,r,:,vphi) = ftemp1(mp)%data(:,r,:)*over_rhor(r)
            Enddo
        Enddo
        ! Velocity_Derivatives
        Call d_by_dtheta(a, d2wdr2, ftemp1)
        Call d_by_dphi(a, dzdr, ftemp2)
        Do mp = 1, h_ang_nm
            ftemp1(mp)%data = ftemp1(mp)%data+ftemp2(mp)%data
            Do r = 1, h_ang_nr
                a(mp)%data(:,r,:,dvtdr) = ftemp1(mp)%data(:,r,:)*over_rhor(r)
            Enddo
            Do r = 1, h_ang_nr
                a(mp)%data(:,r,:,dvtdr) = a(mp)%data(:,r,:,dvtdr)-a(mp)%data(:,r,:,vtheta)*drho_term(r)
            Enddo
        Enddo
        Call d_by_dphi(a, d2wdr2, ftemp1)
        Call d_by_dtheta(a, dzdr, ftemp2)
        Do mp = 1, h_ang_nm
            ftemp1(mp)%data = ftemp1(mp)%data-ftemp2(mp)%data
            Do r = 1, h_ang_nr
                a(mp)%data(:,r,:,dvpdr) = ftemp1(mp)%data(:,r,:)*over_rhor(r)
            Enddo
            Do r = 1, h_ang_nr
                a(mp)%data(:,r,:,dvpdr) = a(mp)%data(:,r,:,dvpdr)-a(mp)%data(:,r,:,vphi)*drho_term(r)
            Enddo
        Enddo
        Do mp = 1, h_ang_nm
            m = h_ang_mvals(mp)
            Do imi = 1, 2
                Do r = 1, h_ang_nr
                    a(mp)%data(m:lm,r,imi,dvrdr) = l_l_plus1(m:lm)*a(mp)%data(m:lm,r,imi,dvrdr)*over_rhorsq(r)
                Enddo
            Enddo
            Do r = 1, h_ang_nr
                a(mp)%data(:,r,:,dvrdr) = a(mp)%data(:,r,:,dvrdr)-a(mp)%data(:,r,:,vr)*two_over_r(r)
            Enddo
            Do r = 1, h_ang_nr
                a(mp)%data(:,r,:,dvrdr) = a(mp)%data(:,r,:,dvrdr)-a(mp)%data(:,r,:,vr)*dlnrho(r)
            Enddo
            If (pseudo /= 0) Then
                Do r = 1, h_ang_nr
                    a(mp)%data(:,r,:,dvrdr) = a(mp)%data(:,r,:,dvrdr)-a(mp)%data(:,r,:,vr)*dsdr_over_cp(r)
                Enddo
            Endif
        Enddo
        Call d_by_dtheta(a, vr, dvrdt)
        Do mp = 1, h_ang_nm
            m = h_ang_mvals(mp)
            Do imi = 1, 2
                Do r = 1, h_ang_nr
                    a(mp)%data(m:lm,r,imi,zvar) = l_l_plus1(m:lm)*a(mp)%data(m:lm,r,imi,zvar)*over_rhorsq(r)
                Enddo
            Enddo
        Enddo
        ! rlm_spacea
        Call d_by_dtheta(a, tvar, dtdt)
        Call pack4d(a, nf, buf)
        DeAllocate(a, ftemp1, ftemp2, l_l_plus1)
    End Subroutine ref_rlm_spacea_hydro

    ! Compute_BandCurlB (Physics/Sphere_Hybrid_Space.F90:377-460) through the reference's
    ! d_by_dtheta / d_by_dphi and the same array statements, in place on a packed buffer of nf
    ! fields.  v = 1-based (C, A, dCdr, dAdr, d2Cdr2, curlbr); Br, Btheta, Bphi, curlbtheta and
    ! curlbphi are the C, A, dCdr, d2Cdr2 and dAdr fields (Fields.F90:329-334).
    Subroutine ref_compute_bandcurlb(nf, v, one_over_rsq, buf) bind(C, name='ref_compute_bandcurlb')
        Integer(C_INT), Value :: nf
        Integer(C_INT), Intent(In) :: v(6)
        Real(C_DOUBLE), Intent(In) :: one_over_rsq(*)
        Real(C_DOUBLE), Intent(InOut) :: buf(*)
        Type(rmcontainer4d), Allocatable :: a(:)
        Type(rmcontainer3d), Allocatable :: ftemp1(:), ftemp2(:)
        Real*8, Allocatable :: l_l_plus1(:)
        Real*8 :: ell
        Integer :: mp, m, l, r, imi, lm
        Integer :: cvar, avar, dcdr, dadr, d2cdr2, curlbr, br, btheta, bphi, curlbtheta, curlbphi
        cvar = v(1); avar = v(2); dcdr = v(3); dadr = v(4); d2cdr2 = v(5); curlbr = v(6)
        br = cvar; btheta = avar; bphi = dcdr; curlbtheta = d2cdr2; curlbphi = dadr
        lm = h_ang_lmax
        Allocate(l_l_plus1(0:lm))
        Do l = 0, lm
            ell = l*1.0d0
            l_l_plus1(l) = ell*(ell+1)
        Enddo
        Call unpack4d(buf, nf, a)
        Call alloc3d(ftemp1)
        Call alloc3d(ftemp2)
        Do mp = 1, h_ang_nm
            m = h_ang_mvals(mp)
            Do imi = 1, 2
                Do r = 1, h_ang_nr
                    a(mp)%data(m:lm,r,imi,br) = l_l_plus1(m:lm)*a(mp)%data(m:lm,r,imi,br)*one_over_rsq(r)
                Enddo
            Enddo
            Do imi = 1, 2
                Do r = 1, h_ang_nr
                    a(mp)%data(m:lm,r,imi,curlbr) = l_l_plus1(m:lm)*a(mp)%data(m:lm,r,imi,avar)*one_over_rsq(r)
                Enddo
            Enddo
            a(mp)%data(:,:,:,d2cdr2) = a(mp)%data(:,:,:,d2cdr2)-a(mp)%data(:,:,:,br)
        Enddo
        Call d_by_dtheta(a, dadr, ftemp1)
        Call d_by_dphi(a, dadr, ftemp2)
        Do mp = 1, h_ang_nm
            a(mp)%data(:,:,:,dadr) = ftemp2(mp)%data
        Enddo
        Call d_by_dtheta(a, d2cdr2, ftemp2)
        Do mp = 1, h_ang_nm
            a(mp)%data(:,:,:,curlbphi) = a(mp)%data(:,:,:,curlbphi)+ftemp2(mp)%data
        Enddo
        Call d_by_dphi(a, d2cdr2, ftemp2)
        Do mp = 1, h_ang_nm
            a(mp)%data(:,:,:,curlbtheta) = ftemp1(mp)%data-ftemp2(mp)%data
        Enddo
        Call d_by_dtheta(a, avar, ftemp1)
        Call d_by_dphi(a, avar, ftemp2)
        Do mp = 1, h_ang_nm
            a(mp)%data(:,:,:,avar) = ftemp2(mp)%data
        Enddo
        Call d_by_dtheta(a, dcdr, ftemp2)
        Do mp = 1, h_ang_nm
            a(mp)%data(:,:,:,avar) = a(mp)%data(:,:,:,avar)+ftemp2(mp)%data
        Enddo
        Call d_by_dphi(a, dcdr, ftemp2)
        Do mp = 1, h_ang_nm
            a(mp)%data(:,:,:,dcdr) = ftemp2(mp)%data-ftemp1(mp)%data
        Enddo
        Call pack4d(a, nf, buf)
        DeAllocate(a, ftemp1, ftemp2, l_l_plus1)
    End Subroutine ref_compute_bandcurlb

    ! Adjust_Emf (Physics/Sphere_Hybrid_Space.F90:557-602) through the reference's d_by_sdtheta /
    ! d_by_dphi and the same array statements, in place on a packed buffer of nf fields; et, ep =
    ! 1-based emftheta (= Cvar, Fields.F90:363-364) and emfphi.
    Subroutine ref_adjust_emf(nf, et, ep, buf) bind(C, name='ref_adjust_emf')
        Integer(C_INT), Value :: nf, et, ep
        Real(C_DOUBLE), Intent(InOut) :: buf(*)
        Type(rmcontainer4d), Allocatable :: a(:)
        Type(rmcontainer3d), Allocatable :: ftemp1(:), ftemp2(:), ftemp3(:)
        Real*8, Allocatable :: ol(:)
        Real*8 :: ell
        Integer :: mp, m, l, r, imi, lm, emftheta, emfphi, cvar
        emftheta = et; emfphi = ep; cvar = et
        lm = h_ang_lmax
        Allocate(ol(0:lm))
        ol(0) = 1.0d0
        Do l = 1, lm
            ell = l*1.0d0
            ol(l) = 1.0d0/(ell*(ell+1))
        Enddo
        Call unpack4d(buf, nf, a)
        Call alloc3d(ftemp1)
        Call alloc3d(ftemp2)
        Call d_by_sdtheta(a, emfphi, ftemp1)
        Call d_by_dphi(a, emftheta, ftemp2)
        Call alloc3d(ftemp3)
        Do mp = 1, h_ang_nm
            ftemp3(mp)%data = a(mp)%data(:,:,:,emftheta)
        Enddo
        Do mp = 1, h_ang_nm
            m = h_ang_mvals(mp)
            Do imi = 1, 2
                Do r = 1, h_ang_nr
                    a(mp)%data(m:lm,r,imi,cvar) = (ftemp1(mp)%data(m:lm,r,imi)-ftemp2(mp)%data(m:lm,r,imi))*ol(m:lm)
                Enddo
            Enddo
        Enddo
        Call d_by_dphi(a, emfphi, ftemp2)
        Do mp = 1, h_ang_nm
            a(mp)%data(:,:,:,emfphi) = ftemp3(mp)%data
        Enddo
        Call d_by_sdtheta(a, emfphi, ftemp1)
        Do mp = 1, h_ang_nm
            m = h_ang_mvals(mp)
            Do imi = 1, 2
                Do r = 1, h_ang_nr
                    a(mp)%data(m:lm,r,imi,emfphi) = (ftemp2(mp)%data(m:lm,r,imi)+ftemp1(mp)%data(m:lm,r,imi))*ol(m:lm)
                Enddo
            Enddo
        Enddo
        Call pack4d(a, nf, buf)
        DeAllocate(a, ftemp1, ftemp2, ftemp3, ol)
    End Subroutine ref_adjust_emf

    ! Hybrid_Output_Initial's magnetic part and Hybrid_Output_Final up to its transform
    ! (Physics/Sphere_Hybrid_Space.F90:224-233,437-552) through the reference's d_by_dtheta /
    ! d_by_dphi (their 3-D and 4-D forms) and the same array statements: s (nf fields), the
    ! co-buffer co (nf_co fields) and the work fields ftemp3, ftemp4 (tmp, two fields) packed;
    ! vi = 1-based (avar, dcdr, d2cdr2, dadr, avar_cb, dbrdr_cb); vf = 1-based (pvar, br,
    ! btheta, bphi, dpdt_cb, dbtdr_cb, dbpdr_cb, dbrdr_cb, dbrdt_cb); mag = 0: the hydro part only.
    Subroutine ref_hybrid_output(nf, nf_co, vi, vf, mag, one_over_r, one_over_rsq, two_over_r, s, co) &
            bind(C, name='ref_hybrid_output')
        Integer(C_INT), Value :: nf, nf_co, mag
        Integer(C_INT), Intent(In) :: vi(6), vf(9)
        Real(C_DOUBLE), Intent(In) :: one_over_r(*), one_over_rsq(*), two_over_r(*)
        Real(C_DOUBLE), Intent(InOut) :: s(*), co(*)
        Type(rmcontainer4d), Allocatable :: a(:), c(:)
        Type(rmcontainer3d), Allocatable :: ftemp1(:), ftemp2(:), ftemp3(:), ftemp4(:)
        Real*8, Allocatable :: l_l_plus1(:)
        Real*8 :: ell
        Integer :: mp, m, l, r, imi, lm
        lm = h_ang_lmax
        Allocate(l_l_plus1(0:lm))
        Do l = 0, lm
            ell = l*1.0d0
            l_l_plus1(l) = ell*(ell+1)
        Enddo
        Call unpack4d(s, nf, a)
        Call unpack4d(co, nf_co, c)
        Call alloc3d(ftemp1)
        Call alloc3d(ftemp2)
        Call alloc3d(ftemp3)
        Call alloc3d(ftemp4)
        ! Hybrid_Output_Initial
        If (mag /= 0) Then
            Do mp = 1, h_ang_nm
                m = h_ang_mvals(mp)
                Do imi = 1, 2
                    Do r = 1, h_ang_nr
                        c(mp)%data(m:lm,r,imi,vi(5)) = l_l_plus1(m:lm)*a(mp)%data(m:lm,r,imi,vi(1))*one_over_rsq(r)
                    Enddo
                Enddo
                c(mp)%data(:,:,:,vi(6)) = a(mp)%data(:,:,:,vi(2))
                ftemp3(mp)%data = a(mp)%data(:,:,:,vi(3))
                ftemp4(mp)%data = a(mp)%data(:,:,:,vi(4))
            Enddo
        Endif
        ! Hybrid_Output_Final
        Do mp = 1, h_ang_nm
            c(mp)%data(lm,:,:,:) = 0.0d0
        Enddo
        Call d_by_dtheta(a, vf(1), ftemp1)
        Do mp = 1, h_ang_nm
            c(mp)%data(:,:,:,vf(5)) = ftemp1(mp)%data
        Enddo
        If (mag /= 0) Then
            Call d_by_dtheta(ftemp3, ftemp1)
            Call d_by_dphi(ftemp4, ftemp2)
            Do mp = 1, h_ang_nm
                ftemp1(mp)%data = ftemp1(mp)%data + ftemp2(mp)%data
                Do r = 1, h_ang_nr
                    c(mp)%data(:,r,:,vf(6)) = ftemp1(mp)%data(:,r,:)*one_over_r(r)
                Enddo
                Do r = 1, h_ang_nr
                    c(mp)%data(:,r,:,vf(6)) = c(mp)%data(:,r,:,vf(6)) - a(mp)%data(:,r,:,vf(3))*one_over_rsq(r)
                Enddo
            Enddo
            Call d_by_dphi(ftemp3, ftemp1)
            Call d_by_dtheta(ftemp4, ftemp2)
            Do mp = 1, h_ang_nm
                m = h_ang_mvals(mp)
                ftemp1(mp)%data = ftemp1(mp)%data - ftemp2(mp)%data
                Do r = 1, h_ang_nr
                    c(mp)%data(:,r,:,vf(7)) = ftemp1(mp)%data(:,r,:)*one_over_r(r)
                Enddo
                Do r = 1, h_ang_nr
                    c(mp)%data(:,r,:,vf(7)) = c(mp)%data(:,r,:,vf(7)) - a(mp)%data(:,r,:,vf(4))*one_over_rsq(r)
                Enddo
                Do imi = 1, 2
                    Do r = 1, h_ang_nr
                        c(mp)%data(m:lm,r,imi,vf(8)) = l_l_plus1(m:lm)*c(mp)%data(m:lm,r,imi,vf(8))*one_over_rsq(r)
                    Enddo
                Enddo
                Do r = 1, h_ang_nr
                    c(mp)%data(:,r,:,vf(8)) = c(mp)%data(:,r,:,vf(8)) - a(mp)%data(:,r,:,vf(2))*two_over_r(r)
                Enddo
            Enddo
            Call d_by_dtheta(a, vf(2), ftemp1)
            Do mp = 1, h_ang_nm
                c(mp)%data(:,:,:,vf(9)) = ftemp1(mp)%data
            Enddo
        Endif
        Call pack4d(c, nf_co, co)
        DeAllocate(a, c, ftemp1, ftemp2, ftemp3, ftemp4, l_l_plus1)
    End Subroutine ref_hybrid_output

End Module raysht_ref_harness
