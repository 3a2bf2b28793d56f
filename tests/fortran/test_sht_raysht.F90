!
! Test program for the RaySHT Fortran binding, written the way Rayleigh's own
! Test_SHT (src/Test_Suite/Test_SHT.F90) exercises its transforms, on one rank:
! single (l,m) = (2,1) mode against the analytic field, 10 white-noise round trips,
! the Legendre + transpose + FFT loop (index loops, then Reform_2a3a / Reform_3b2b, then
! the device-resident procedures), the radial reforms, the 2-D Legendre pair, the Real*16
! grid, and the Chebyshev transforms both through Cheby_Init and through a Cheby_Grid
! object called as Rayleigh calls gridcp (Test_Cheby.F90).  Prints one PASSING / FAIL line per check (tolerance 1e-10, as Test_SHT.F90:34)
! and exits non-zero on any failure.
!
Program test_sht_raysht
    Use ISO_C_BINDING
    Use Structures
    Use RaySHT_Math_Layer
    Implicit None
    Integer, Parameter :: lmax = 41, nt = 64, nphi = 128, nr = 4
    Real*8, Parameter :: tol = 1.0d-10, pi = 3.1415926535897932384626433832795d0
    Integer :: mvals(lmax+1), nfail
    Integer :: i

    ! balanced one-rank m order (Load_Balance.F90:126-226): 0, L, 1, L-1, ...
    Do i = 0, (lmax+1)/2 - 1
        mvals(2*i+1) = i
        mvals(2*i+2) = lmax - i
    Enddo
    nfail = 0
    Call Initialize_Legendre(nt, lmax, mvals, .true.)
    Call single_mode(nfail)
    Call white_noise(nfail)
    Call lt_fft_loop(nfail)
    Call reform_loop(nfail)
    Call radial_reforms(nfail)
    Call real16_grid(nfail)
    Call two_d_pair(nfail)
    Call cheby_round_trip(nfail)
    Call cheby_derivatives(nfail)
    Call angular_derivatives(nfail)
    Call implicit_solves(nfail)
    Call cheby_grid_object(nfail)
    Call spectral_passes(nfail)
    Call output_final_hydro(nfail)
    ! the split forms with 6 digits stay within the same tolerance over ten round trips
    Call Set_Legendre_Precision(raysht_int8x6)
    Call white_noise(nfail)
    Call Set_Legendre_Precision(raysht_bf16x6)
    Call white_noise(nfail)
    Call Set_Legendre_Precision(raysht_fp64)
    Call Finalize_Legendre()
    If (nfail > 0) Then
        Write(6, '(A, I0, A)') 'RaySHT Fortran tests: ', nfail, ' FAILED'
        Error Stop 1
    Endif
    Write(6, '(A)') 'RaySHT Fortran tests: all PASSING'

Contains

    Subroutine report(name, err, nfail)
        Character(len=*), Intent(In) :: name
        Real*8, Intent(In) :: err
        Integer, Intent(InOut) :: nfail
        If (err < tol) Then
            Write(6, '(A, A, ES10.3)') name, ' result = PASSING, max error ', err
        Else
            Write(6, '(A, A, ES10.3)') name, ' result = FAIL, max error ', err
            nfail = nfail + 1
        Endif
    End Subroutine report

    Subroutine alloc_spec(s, nf)
        Type(rmcontainer4d), Allocatable, Intent(InOut) :: s(:)
        Integer, Intent(In) :: nf
        Integer :: mp
        Allocate(s(lmax+1))
        Do mp = 1, lmax+1
            Allocate(s(mp)%data(mvals(mp):lmax, 1:nr, 1:2, 1:nf))
            s(mp)%data = 0.0d0
        Enddo
    End Subroutine alloc_spec

    ! Test_SHT.F90:47-200
    Subroutine single_mode(nfail)
        Integer, Intent(InOut) :: nfail
        Integer, Parameter :: nf = 3
        Type(rmcontainer4d), Allocatable :: s(:)
        Real*8, Allocatable :: p(:,:,:), truth(:)
        Real*8 :: radius(nr), d_phys, d_spec, amp, norm
        Integer :: mp, f, r, ind
        Call alloc_spec(s, nf)
        Allocate(p(nt, 2*nr*nf, lmax+1), truth(nt))
        Do r = 1, nr
            radius(r) = 0.5d0 + (r-1)*1.0d0/(nr-1)
        Enddo
        truth(:) = sqrt(5.0d0/(4.0d0*pi*6.0d0))*(-3.0d0*coloc(:)*sqrt(1.0d0-coloc(:)**2))
        Do mp = 1, lmax+1
            If (mvals(mp) /= 1) Cycle
            Do f = 1, nf
                s(mp)%data(2, :, 1, f) = (2.0d0*f+1.0d0)*radius(:)*2.0d0
                s(mp)%data(2, :, 2, f) = -(3.0d0*f*f+2.0d0)*radius(:)**2*2.0d0
            Enddo
        Enddo
        Call Legendre_Transform(s, p)
        d_phys = 0.0d0
        Do mp = 1, lmax+1
            If (mvals(mp) == 1) Then
                Do f = 1, nf
                    Do r = 1, nr
                        ind = r + (f-1)*nr*2
                        amp = (2.0d0*f+1.0d0)*radius(r)
                        d_phys = max(d_phys, maxval(abs(amp*truth(:) - p(:, ind, mp))))
                        ind = r + nr + (f-1)*nr*2
                        amp = -(3.0d0*f*f+2.0d0)*radius(r)**2
                        d_phys = max(d_phys, maxval(abs(amp*truth(:) - p(:, ind, mp))))
                    Enddo
                Enddo
            Else
                d_phys = max(d_phys, maxval(abs(p(:, :, mp))))
            Endif
        Enddo
        Call report('Spec -> Phys', d_phys, nfail)
        Call Legendre_Transform(p, s)
        d_spec = 0.0d0
        Do mp = 1, lmax+1
            norm = nt
            s(mp)%data(lmax, :, :, :) = 0.0d0
            s(mp)%data = s(mp)%data*norm
            If (mvals(mp) == 1) Then
                Do f = 1, nf
                    d_spec = max(d_spec, maxval(abs(s(mp)%data(2, :, 1, f) - (2.0d0*f+1.0d0)*radius(:))))
                    d_spec = max(d_spec, maxval(abs(s(mp)%data(2, :, 2, f) + (3.0d0*f*f+2.0d0)*radius(:)**2)))
                    s(mp)%data(2, :, :, f) = 0.0d0
                Enddo
            Endif
            d_spec = max(d_spec, maxval(abs(s(mp)%data)))
        Enddo
        Call report('Spec -> Phys -> Spec', d_spec, nfail)
    End Subroutine single_mode

    ! Test_SHT.F90:328-447
    Subroutine white_noise(nfail)
        Integer, Intent(InOut) :: nfail
        Integer, Parameter :: nf = 3, nloops = 10
        Type(rmcontainer4d), Allocatable :: s(:)
        Real*8, Allocatable :: p(:,:,:)
        Real*8 :: d, to_phys, to_spec
        Integer :: mp, m, it
        Call alloc_spec(s, nf)
        Allocate(p(nt, 2*nr*nf, lmax+1))
        Do mp = 1, lmax+1
            m = mvals(mp)
            If (m < lmax) s(mp)%data(m:lmax-1, :, :, :) = 1.0d0
            If (m == 0) s(mp)%data(:, :, 2, :) = 0.0d0
        Enddo
        Do it = 1, nloops
            Do mp = 1, lmax+1
                to_phys = 0.5d0
                If (mvals(mp) == 0) to_phys = 1.0d0
                s(mp)%data = s(mp)%data/to_phys
            Enddo
            Call Legendre_Transform(s, p)
            Call Legendre_Transform(p, s)
            Do mp = 1, lmax+1
                to_spec = 1.0d0/nt
                If (mvals(mp) == 0) to_spec = 1.0d0/nphi
                s(mp)%data(lmax, :, :, :) = 0.0d0
                s(mp)%data = s(mp)%data/to_spec
            Enddo
        Enddo
        d = 0.0d0
        Do mp = 1, lmax+1
            m = mvals(mp)
            If (m < lmax) Then
                If (m > 0) Then
                    d = max(d, maxval(abs(1.0d0 - s(mp)%data(m:lmax-1, :, :, :))))
                Else
                    d = max(d, maxval(abs(1.0d0 - s(mp)%data(m:lmax-1, :, 1, :))))
                    d = max(d, maxval(abs(s(mp)%data(m:lmax-1, :, 2, :))))
                Endif
            Endif
            d = max(d, maxval(abs(s(mp)%data(lmax, :, :, :))))
        Enddo
        Call report('Spec -> Phys -> Spec -> ... (10 loops)', d/nloops, nfail)
    End Subroutine white_noise

    ! Test_SHT.F90:449-575, the one-rank transposes written out as index maps
    Subroutine lt_fft_loop(nfail)
        Integer, Intent(InOut) :: nfail
        Integer, Parameter :: nf = 5
        Type(rmcontainer4d), Allocatable :: s(:)
        Real*8, Allocatable :: p(:,:,:), x(:,:,:,:), norm(:,:)
        Real*8 :: d, ans
        Integer :: mp, m, l, f, r, t
        Call alloc_spec(s, nf)
        Allocate(p(nt, 2*nr*nf, lmax+1), x(nphi+2, nr, nt, nf), norm(lmax+1, 2))
        Do mp = 1, lmax+1
            m = mvals(mp)
            Do l = m, lmax-1
                Do r = 1, nr
                    Do f = 1, nf
                        s(mp)%data(l, r, 1, f) = (l+1)*1.0d0*(2*f+3)*r
                        s(mp)%data(l, r, 2, f) = (m+1)*1.0d0*(4*f+5)*r*r
                    Enddo
                Enddo
            Enddo
            norm(mp, 1) = max(maxval(abs(s(mp)%data(:, :, 1, :))), 1.0d-300)
            norm(mp, 2) = max(maxval(abs(s(mp)%data(:, :, 2, :))), 1.0d-300)
            s(mp)%data(:, :, 1, :) = s(mp)%data(:, :, 1, :)/norm(mp, 1)
            s(mp)%data(:, :, 2, :) = s(mp)%data(:, :, 2, :)/norm(mp, 2)
            s(mp)%data(lmax, :, :, :) = 0.0d0
            If (m == 0) s(mp)%data(:, :, 2, :) = 0.0d0
        Enddo
        Call Legendre_Transform(s, p)
        x(:, :, :, :) = 0.0d0
        Do mp = 1, lmax+1
            m = mvals(mp)
            Do f = 1, nf
                Do t = 1, nt
                    Do r = 1, nr
                        x(2*m+1, r, t, f) = p(t, r + (f-1)*2*nr, mp)
                        x(2*m+2, r, t, f) = p(t, r + nr + (f-1)*2*nr, mp)
                    Enddo
                Enddo
            Enddo
        Enddo
        Call FFT_To_Physical(x, rsc=.true.)
        Call FFT_To_Spectral(x, rsc=.true.)
        Do mp = 1, lmax+1
            m = mvals(mp)
            Do f = 1, nf
                Do t = 1, nt
                    Do r = 1, nr
                        p(t, r + (f-1)*2*nr, mp) = x(2*m+1, r, t, f)
                        p(t, r + nr + (f-1)*2*nr, mp) = x(2*m+2, r, t, f)
                    Enddo
                Enddo
            Enddo
        Enddo
        Call Legendre_Transform(p, s)
        d = 0.0d0
        Do mp = 1, lmax+1
            m = mvals(mp)
            Do l = m, lmax-1
                Do r = 1, nr
                    Do f = 1, nf
                        ans = (l+1)*1.0d0*(2*f+3)*r/norm(mp, 1)
                        d = max(d, abs(ans - s(mp)%data(l, r, 1, f)))
                        ans = 0.0d0
                        If (m > 0) ans = (m+1)*1.0d0*(4*f+5)*r*r/norm(mp, 2)
                        d = max(d, abs(ans - s(mp)%data(l, r, 2, f)))
                    Enddo
                Enddo
            Enddo
        Enddo
        Call report('LT/FFT/transpose', d, nfail)
    End Subroutine lt_fft_loop

    ! Test_SHT.F90:449-575 with SphericalBuffer%reform's one-rank transposes as Reform_2a3a /
    ! Reform_3b2b (checked against the index loops above), then the same chain on device
    ! buffers through the *_Device procedures (checked bit for bit against the host arrays)
    Subroutine reform_loop(nfail)
        Integer, Intent(InOut) :: nfail
        Integer, Parameter :: nf = 5
        Type(rmcontainer4d), Allocatable :: s(:)
        Real*8, Allocatable, Target :: p(:,:,:), x(:,:,:,:), xi(:,:,:,:), norm(:,:), packed(:), back(:)
        Type(C_PTR) :: ds, dp, dx
        Real*8 :: d, dd, ans
        Integer :: mp, m, l, f, r, t, off, nl
        Integer*8 :: nbs, nbp, nbx
        Call alloc_spec(s, nf)
        Allocate(p(nt, 2*nr*nf, lmax+1), x(nphi+2, nr, nt, nf), xi(nphi+2, nr, nt, nf), norm(lmax+1, 2))
        Do mp = 1, lmax+1
            m = mvals(mp)
            Do l = m, lmax-1
                Do r = 1, nr
                    Do f = 1, nf
                        s(mp)%data(l, r, 1, f) = (l+1)*1.0d0*(2*f+3)*r
                        s(mp)%data(l, r, 2, f) = (m+1)*1.0d0*(4*f+5)*r*r
                    Enddo
                Enddo
            Enddo
            norm(mp, 1) = max(maxval(abs(s(mp)%data(:, :, 1, :))), 1.0d-300)
            norm(mp, 2) = max(maxval(abs(s(mp)%data(:, :, 2, :))), 1.0d-300)
            s(mp)%data(:, :, 1, :) = s(mp)%data(:, :, 1, :)/norm(mp, 1)
            s(mp)%data(:, :, 2, :) = s(mp)%data(:, :, 2, :)/norm(mp, 2)
            s(mp)%data(lmax, :, :, :) = 0.0d0
            If (m == 0) s(mp)%data(:, :, 2, :) = 0.0d0
        Enddo
        ! the device copy of the input: blocks packed in m_values order
        nbs = 0
        Do mp = 1, lmax+1
            nbs = nbs + size(s(mp)%data)
        Enddo
        Allocate(packed(nbs), back(nbs))
        off = 0
        Do mp = 1, lmax+1
            nl = size(s(mp)%data)
            packed(off+1:off+nl) = reshape(s(mp)%data, (/ nl /))
            off = off + nl
        Enddo
        nbs = nbs * 8
        nbp = int(size(p), 8) * 8
        nbx = int(size(x), 8) * 8
        ds = Device_Allocate(nbs)
        dp = Device_Allocate(nbp)
        dx = Device_Allocate(nbx)
        Call Device_Upload(ds, c_loc(packed), nbs)
        ! host arrays: Legendre_Transform, Reform_2a3a, the FFTs, Reform_3b2b, Legendre_Transform
        Call Legendre_Transform(s, p)
        Call Reform_2a3a(p, x)
        xi(:, :, :, :) = 0.0d0
        Do mp = 1, lmax+1
            m = mvals(mp)
            Do f = 1, nf
                Do t = 1, nt
                    Do r = 1, nr
                        xi(2*m+1, r, t, f) = p(t, r + (f-1)*2*nr, mp)
                        xi(2*m+2, r, t, f) = p(t, r + nr + (f-1)*2*nr, mp)
                    Enddo
                Enddo
            Enddo
        Enddo
        dd = maxval(abs(x - xi))          ! a pure permutation: exactly the index loop
        Call FFT_To_Physical(x, rsc=.true.)
        Call FFT_To_Spectral(x, rsc=.true.)
        Call Reform_3b2b(x, p)
        Call Legendre_Transform(p, s)
        d = 0.0d0
        Do mp = 1, lmax+1
            m = mvals(mp)
            Do l = m, lmax-1
                Do r = 1, nr
                    Do f = 1, nf
                        ans = (l+1)*1.0d0*(2*f+3)*r/norm(mp, 1)
                        d = max(d, abs(ans - s(mp)%data(l, r, 1, f)))
                        ans = 0.0d0
                        If (m > 0) ans = (m+1)*1.0d0*(4*f+5)*r*r/norm(mp, 2)
                        d = max(d, abs(ans - s(mp)%data(l, r, 2, f)))
                    Enddo
                Enddo
            Enddo
        Enddo
        Call report('LT/FFT/reform (Reform_2a3a, Reform_3b2b)', max(d, dd), nfail)
        ! device buffers: the same six steps, nothing copied between them
        Call StP_Device(nr, nf, ds, dp)
        Call Reform_2a3a_Device(nr, nf, nphi, dp, dx)
        Call FFT_To_Physical_Device(nphi, int(nr, 8)*nt*nf, dx)
        Call FFT_To_Spectral_Device(nphi, int(nr, 8)*nt*nf, dx)
        Call Reform_3b2b_Device(nr, nf, nphi, dx, dp)
        Call PtS_Device(nr, nf, dp, ds)
        Call Device_Synchronize()
        Call Device_Download(c_loc(back), ds, nbs)
        Call Device_Free(ds)
        Call Device_Free(dp)
        Call Device_Free(dx)
        d = 0.0d0
        off = 0
        Do mp = 1, lmax+1
            nl = size(s(mp)%data)
            d = max(d, maxval(abs(back(off+1:off+nl) - reshape(s(mp)%data, (/ nl /)))))
            off = off + nl
        Enddo
        Call report('LT/FFT/reform on device buffers (equal to host arrays)', d, nfail)
        If (d /= 0.0d0) Then
            Write(6, '(A)') 'device path differs from the host-array path'
            nfail = nfail + 1
        Endif
    End Subroutine reform_loop

    ! Transpose_2b1b then Transpose_1a2a (Spherical_Buffer.F90:705-932) as Reform_2b1b /
    ! Reform_1a2a: a permutation of the values (same sum of squares) whose inverse returns
    ! the spectral array exactly
    Subroutine radial_reforms(nfail)
        Integer, Intent(InOut) :: nfail
        Integer, Parameter :: nf = 3
        Type(rmcontainer4d), Allocatable :: s(:), s2(:)
        Real*8, Allocatable :: p1(:,:,:,:)
        Real*8 :: d, ss, sp
        Integer :: mp, m, l
        Call alloc_spec(s, nf)
        Call alloc_spec(s2, nf)
        ss = 0.0d0
        Do mp = 1, lmax+1
            m = mvals(mp)
            Do l = m, lmax
                s(mp)%data(l, :, 1, :) = sin(0.3d0*l + 0.17d0*m)
                s(mp)%data(l, :, 2, :) = cos(0.7d0*l - 0.29d0*m)
            Enddo
            ss = ss + sum(s(mp)%data**2)
        Enddo
        Allocate(p1(nr, 2, LM_Count(nr, nf), nf))
        Call Reform_2b1b(s, p1)
        sp = sum(p1**2)
        Call Reform_1a2a(p1, s2)
        d = abs(ss - sp) / ss
        Do mp = 1, lmax+1
            d = max(d, maxval(abs(s2(mp)%data - s(mp)%data)))
        Enddo
        Call report('Reform_2b1b / Reform_1a2a', d, nfail)
    End Subroutine radial_reforms

    ! coloc and gl_weights as the reference keeps them (Real*16): ascending from -1,
    ! symmetric, weights summing to 2
    Subroutine real16_grid(nfail)
        Integer, Intent(InOut) :: nfail
        Real*8 :: d
        Integer :: i
        d = real(abs(sum(gl_weights) - 2.0_16), 8)
        Do i = 1, nt
            d = max(d, real(abs(coloc(i) + coloc(nt+1-i)), 8))
            If (i > 1) Then
                If (coloc(i) <= coloc(i-1)) d = 1.0d0
            Endif
        Enddo
        If (kind(coloc) /= 16) d = 1.0d0
        Call report('Real*16 Gauss-Legendre grid', d, nfail)
    End Subroutine real16_grid

    ! StP_2d_dgp / PtS_2d_dgpv2 with 7 anonymous columns: the round trip is the identity / N_phi
    Subroutine two_d_pair(nfail)
        Integer, Intent(InOut) :: nfail
        Integer, Parameter :: ncol = 7
        Type(rmcontainer), Allocatable :: s(:), s0(:)
        Real*8, Allocatable :: p(:,:,:)
        Real*8 :: d
        Integer :: mp, m, l, c
        Allocate(s(lmax+1), s0(lmax+1), p(nt, ncol, lmax+1))
        Do mp = 1, lmax+1
            m = mvals(mp)
            Allocate(s(mp)%data(m:lmax, 1:ncol))
            Do c = 1, ncol
                Do l = m, lmax
                    s(mp)%data(l, c) = cos(0.37d0*l + 1.3d0*c + 0.11d0*m)/(l+1)
                Enddo
            Enddo
            s0(mp) = s(mp)
        Enddo
        Call Legendre_Transform(s, p)
        Call Legendre_Transform(p, s)
        d = 0.0d0
        Do mp = 1, lmax+1
            d = max(d, maxval(abs(s(mp)%data*nphi - s0(mp)%data)))
        Enddo
        Call report('2-D Legendre pair', d, nfail)
    End Subroutine two_d_pair

    ! To_Spectral4D then From_Spectral4D (Chebyshev_Polynomials.F90:375-558)
    Subroutine cheby_round_trip(nfail)
        Integer, Intent(InOut) :: nfail
        Integer, Parameter :: n = 48
        Real*8 :: grid(n), f(n, 2, 3, 2), c(n, 2, 3, 2), g(n, 2, 3, 2), d
        Integer :: i
        Call Cheby_Init(n, 0.35d0, 1.0d0, grid)
        Do i = 1, n
            f(i, :, :, :) = exp(-grid(i))
            f(i, 2, :, :) = grid(i)**3
        Enddo
        Call Cheby_To_Spectral(f, c)
        Call Cheby_From_Spectral(c, g)
        d = maxval(abs(g - f))
        Call Cheby_Finalize()
        Call report('Chebyshev round trip', d, nfail)
    End Subroutine cheby_round_trip

    ! Spectral_Derivatives.F90 through the generics: field 1 = Y_20 (m = 0, real) and Y_11
    ! (m = 1, real); d_by_dtheta into field 2 and d_by_sdtheta into a separate array against
    ! sin(t) dY_lm/dt = l c_{l+1,m} Y_{l+1,m} - (l+1) c_lm Y_{l-1,m} (c_lm = sqrt((l^2-m^2)/(4l^2-1)))
    ! and 1/sin(t) d(sin^2(t) Y)/dt = (l+2) c_{l+1,m} Y_{l+1,m} - (l-1) c_lm Y_{l-1,m};
    ! d_by_dphi in place: i m Y_11 -> imaginary part m
    Subroutine angular_derivatives(nfail)
        Integer, Intent(InOut) :: nfail
        Type(rmcontainer4d), Allocatable :: s(:)
        Type(rmcontainer3d), Allocatable :: arr(:)
        Real*8 :: d, c20, c30, c21, want(0:lmax)
        Integer :: mp, m0, m1
        Call alloc_spec(s, 2)
        Allocate(arr(lmax+1))
        Do mp = 1, lmax+1
            Allocate(arr(mp)%data(mvals(mp):lmax, 1:nr, 1:2))
            arr(mp)%data = 0.0d0
            If (mvals(mp) == 0) m0 = mp
            If (mvals(mp) == 1) m1 = mp
        Enddo
        s(m0)%data(2, :, 1, 1) = 1.0d0
        s(m1)%data(1, :, 1, 1) = 1.0d0
        c20 = sqrt(4.0d0 / 15.0d0)   ! c_{2,0}
        c30 = sqrt(9.0d0 / 35.0d0)   ! c_{3,0}
        c21 = sqrt(3.0d0 / 15.0d0)   ! c_{2,1}
        Call d_by_dtheta(s, 1, 2)
        want = 0.0d0
        want(1) = -3.0d0 * c20
        want(3) = 2.0d0 * c30
        d = maxval(abs(s(m0)%data(0:lmax, 1, 1, 2) - want(0:lmax)))
        want = 0.0d0
        want(2) = 1.0d0 * c21        ! l = 1, m = 1: l c_{2,1} Y_21 (no l - 1 = 0 term)
        d = max(d, maxval(abs(s(m1)%data(1:lmax, 2, 1, 2) - want(1:lmax))))
        Call d_by_sdtheta(s, 1, arr)
        want = 0.0d0
        want(1) = -1.0d0 * c20
        want(3) = 4.0d0 * c30
        d = max(d, maxval(abs(arr(m0)%data(0:lmax, 3, 1) - want(0:lmax))))
        Call d_by_dphi(s, 1, 1)
        d = max(d, abs(s(m1)%data(1, 4, 1, 1)) + abs(s(m1)%data(1, 4, 2, 1) - 1.0d0))
        Call report('Angular derivatives', d, nfail)
    End Subroutine angular_derivatives

    ! Implicit_Solve's two paths as Linear_Solve.F90 calls them: LU_Decompose_full / LU_Solve_full
    ! on a 12 x 12 matrix and LU_Decompose_band / LU_Solve_Band on a 20 x 20 matrix of band width
    ! kl = ku = 2 in band storage (ldab = 7, element (r, c) at row 5 + r - c, as Band_Load_Single
    ! lays it out), each with 2 x 2 right-hand sides A x for a known x
    Subroutine implicit_solves(nfail)
        Integer, Intent(InOut) :: nfail
        Integer, Parameter :: nd = 12, nb = 20, kl = 2
        Real*8 :: a(nd, nd), x(nd, 2, 2), rhs(nd, 2, 2), band(3*kl+1, nb), full(nb, nb)
        Real*8 :: xb(nb, 2, 2), rb(nb, 2, 2), d
        Integer :: pvt(nd), pb(nb), r, c, i, j
        Do c = 1, nd
            Do r = 1, nd
                a(r, c) = sin(1.3d0 * r + 0.7d0 * c * c)
            Enddo
            a(c, c) = a(c, c) + 0.1d0 * c
        Enddo
        Do j = 1, 2
            Do i = 1, 2
                Do r = 1, nd
                    x(r, i, j) = cos(0.5d0 * r + i - 2.0d0 * j)
                Enddo
                rhs(:, i, j) = matmul(a, x(:, i, j))
            Enddo
        Enddo
        Call LU_Decompose_full(a, pvt)
        Call LU_Solve_full(a, rhs, pvt)
        d = maxval(abs(rhs - x))
        full = 0.0d0
        band = 0.0d0
        Do c = 1, nb
            Do r = max(1, c - kl), min(nb, c + kl)
                full(r, c) = cos(0.9d0 * r - 0.4d0 * c) + merge(0.0d0, 0.2d0, r /= c)
                band(2*kl + 1 + r - c, c) = full(r, c)
            Enddo
        Enddo
        Do j = 1, 2
            Do i = 1, 2
                Do r = 1, nb
                    xb(r, i, j) = sin(0.3d0 * r * i + j)
                Enddo
                rb(:, i, j) = matmul(full, xb(:, i, j))
            Enddo
        Enddo
        Call LU_Decompose_band(band, pb)
        Call LU_Solve_Band(band, rb, pb)
        d = max(d, maxval(abs(rb - xb)))
        Call report('Implicit solves (dense, band)', d, nfail)
    End Subroutine implicit_solves

    ! A Cheby_Grid object used as Rayleigh uses gridcp (ProblemSize.F90:539, Test_Cheby.F90:
    ! 230-340): Init with the grid and integration weights, to_spectral / from_spectral of a
    ! 4-D array, d_by_dr_cp into other variables of the buffer, d_by_dr_nobuffer, and
    ! dealias_buffer; the grid equals Cheby_Init's, the weights integrate r^2 dr over the shell
    ! to 1 (to the quadrature's accuracy)
    Subroutine cheby_grid_object(nfail)
        Integer, Intent(InOut) :: nfail
        Integer, Parameter :: n = 48, nd = 8
        Type(Cheby_Grid) :: gridcp, small
        Real*8 :: radius(n), rw(n), g2(n), f(n, 3, 3, 2), c(n, 3, 3, 2), fc(n, 3, 3, 2), c2(n, 3, 3, 2)
        Real*8 :: r8(nd), w8(nd), fd(nd, 2, 2, 4), cd(nd, 2, 2, 4), ans(nd, 2, 2, 4), ob(nd, 2, 2), d
        Integer :: npoly(1), i, j
        npoly(1) = n
        Call gridcp%Init(radius, rw, 1, npoly, (/ 0.35d0, 1.0d0 /), nthread = 1, dealias_by = (/ -1 /))
        Call Cheby_Init(n, 0.35d0, 1.0d0, g2)
        Call Cheby_Finalize()
        d = maxval(abs(radius - g2))
        If (abs(sum(rw) - 1.0d0) > 1.0d-3 .or. gridcp%npoly(1) /= n .or. gridcp%rda(1) /= (2*n)/3 + 1) d = 1.0d0
        Do j = 1, 3
            Do i = 1, 3
                f(:, i, j, 1) = exp(-4.0d0*radius(:)**2) + exp(-j*radius(:)**2)
            Enddo
        Enddo
        f(:, :, :, 2) = 2.0d0*f(:, :, :, 1)
        Call gridcp%to_spectral(f, c)
        Call gridcp%from_spectral(c, fc)
        d = max(d, maxval(abs(fc - f)))
        c2 = c
        Call gridcp%dealias_buffer(c2)
        d = max(d, maxval(abs(c2(gridcp%rda(1):n, :, :, :))) + maxval(abs(c2(1:gridcp%rda(1)-1, :, :, :) - &
                                                                           c(1:gridcp%rda(1)-1, :, :, :))))
        Call gridcp%finalize()
        ! derivatives of r^3 on [0.5, 2] (exact: 3 r^2, 6 r, 6)
        npoly(1) = nd
        Call small%Init(r8, w8, 1, npoly, (/ 0.5d0, 2.0d0 /))
        fd = 0.0d0
        Do i = 1, nd
            fd(i, :, :, 1) = r8(i)**3
            ans(i, :, :, 2) = 3.0d0 * r8(i)**2
            ans(i, :, :, 3) = 6.0d0 * r8(i)
            ans(i, :, :, 4) = 6.0d0
        Enddo
        Call small%to_spectral(fd, cd)
        Call small%d_by_dr_cp(1, 2, cd, 1)
        Call small%d_by_dr_cp(1, 3, cd, 2)
        Call small%d_by_dr_cp(1, 4, cd, 3)
        Call small%d_by_dr_nobuffer(1, cd, ob, 1)
        d = max(d, maxval(abs(ob - cd(:, :, :, 2))))
        Call small%from_spectral(cd, fd)
        d = max(d, maxval(abs(fd(:, :, :, 2:4) - ans(:, :, :, 2:4))))
        Call small%finalize()
        Call report('Cheby_Grid object (Init, to/from_spectral, d_by_dr, dealias)', d, nfail)
    End Subroutine cheby_grid_object

    ! Cheby_Deriv_Buffer_4D (d_by_dr_cp, Chebyshev_Polynomials.F90:560-640) as Test_Cheby's
    ! derivative check uses it: variables 2..4 of the buffer receive the first three
    ! derivatives of variable 1 = r^3 on [0.5, 2] (exact: 3 r^2, 6 r, 6)
    Subroutine cheby_derivatives(nfail)
        Integer, Intent(InOut) :: nfail
        Integer, Parameter :: n = 8
        Real*8 :: grid(n), f(n, 2, 3, 4), c(n, 2, 3, 4), ans(n, 2, 3, 4), d
        Integer :: i
        Call Cheby_Init(n, 0.5d0, 2.0d0, grid)
        f = 0.0d0
        Do i = 1, n
            f(i, :, :, 1) = grid(i)**3
            ans(i, :, :, 2) = 3.0d0 * grid(i)**2
            ans(i, :, :, 3) = 6.0d0 * grid(i)
            ans(i, :, :, 4) = 6.0d0
        Enddo
        Call Cheby_To_Spectral(f, c)
        Call Cheby_d_by_dr_cp(1, 2, c, 1)
        Call Cheby_d_by_dr_cp(1, 3, c, 2)
        Call Cheby_d_by_dr_cp(1, 4, c, 3)
        Call Cheby_From_Spectral(c, f)
        d = maxval(abs(f(:, :, :, 2:4) - ans(:, :, :, 2:4)))
        Call Cheby_Finalize()
        Call report('Chebyshev derivatives', d, nfail)
    End Subroutine cheby_derivatives

    ! Adjust_Emf_Device (Sphere_Hybrid_Space.F90:557-602) on a device buffer against the
    ! routine's own statements over the module's host d_by_sdtheta / d_by_dphi (the same
    ! doubles), then Zero_Lmax_Device
    Subroutine spectral_passes(nfail)
        Integer, Intent(InOut) :: nfail
        Integer, Parameter :: nf = 3, et = 1, ep = 3
        Type(rmcontainer4d), Allocatable :: s(:)
        Type(rmcontainer3d), Allocatable :: t1(:), t2(:), t3(:)
        Real*8, Allocatable, Target :: packed(:), back(:)
        Type(C_PTR) :: dbuf
        Integer :: mp, m, l, r, imi, f, n, off, total
        Integer*8 :: nb
        Real*8 :: d, ol(0:lmax), ell
        Call alloc_spec(s, nf)
        Allocate(t1(lmax+1), t2(lmax+1), t3(lmax+1))
        total = 0
        Do mp = 1, lmax+1
            m = mvals(mp)
            Allocate(t1(mp)%data(m:lmax, 1:nr, 1:2), t2(mp)%data(m:lmax, 1:nr, 1:2), t3(mp)%data(m:lmax, 1:nr, 1:2))
            Do f = 1, nf
                Do imi = 1, 2
                    Do r = 1, nr
                        Do l = m, lmax
                            s(mp)%data(l, r, imi, f) = sin(0.37d0*l + 1.1d0*r + 0.5d0*imi + 2.3d0*f + 0.013d0*m)
                        Enddo
                    Enddo
                Enddo
            Enddo
            total = total + size(s(mp)%data)
        Enddo
        ol(0) = 1.0d0
        Do l = 1, lmax
            ell = l*1.0d0
            ol(l) = 1.0d0/(ell*(ell+1))
        Enddo
        Allocate(packed(total), back(total))
        Call pack_spec(s, packed)
        nb = int(total, 8) * 8
        dbuf = Device_Allocate(nb)
        Call Device_Upload(dbuf, c_loc(packed), nb)
        Call Adjust_Emf_Device(nr, nf, dbuf, et, ep)
        Call Device_Download(c_loc(back), dbuf, nb)
        ! the routine's statements on the host arrays
        Call d_by_sdtheta(s, ep, t1)
        Call d_by_dphi(s, et, t2)
        Do mp = 1, lmax+1
            m = mvals(mp)
            t3(mp)%data = s(mp)%data(:, :, :, et)
            Do imi = 1, 2
                Do r = 1, nr
                    s(mp)%data(m:lmax, r, imi, et) = (t1(mp)%data(m:lmax, r, imi) - t2(mp)%data(m:lmax, r, imi))*ol(m:lmax)
                Enddo
            Enddo
        Enddo
        Call d_by_dphi(s, ep, t2)
        Do mp = 1, lmax+1
            s(mp)%data(:, :, :, ep) = t3(mp)%data
        Enddo
        Call d_by_sdtheta(s, ep, t1)
        Do mp = 1, lmax+1
            m = mvals(mp)
            Do imi = 1, 2
                Do r = 1, nr
                    s(mp)%data(m:lmax, r, imi, ep) = (t2(mp)%data(m:lmax, r, imi) + t1(mp)%data(m:lmax, r, imi))*ol(m:lmax)
                Enddo
            Enddo
        Enddo
        Call pack_spec(s, packed)
        d = maxval(abs(packed - back))
        Call Zero_Lmax_Device(nr, nf, dbuf)
        Call Device_Download(c_loc(back), dbuf, nb)
        off = 0
        Do mp = 1, lmax+1
            n = lmax - mvals(mp) + 1
            Do f = 1, 2 * nr * nf
                d = max(d, abs(back(off + f * n)))
            Enddo
            off = off + n * 2 * nr * nf
        Enddo
        Call Device_Free(dbuf)
        Call report('Adjust_Emf_Device / Zero_Lmax_Device', d, nfail)
    End Subroutine spectral_passes

    ! Hybrid_Output_Final_Device without magnetism against its statements on the host arrays:
    ! every co field's l_max zeroed, then co(dpdt_cb) = d_by_dtheta s(pvar)
    Subroutine output_final_hydro(nfail)
        Integer, Intent(InOut) :: nfail
        Integer, Parameter :: nf = 3, pvar = 2, dpdt_cb = 3
        Type(rmcontainer4d), Allocatable :: s(:), co(:)
        Type(rmcontainer3d), Allocatable :: t1(:)
        Real*8, Allocatable, Target :: packed(:), back(:)
        Type(C_PTR) :: ds, dco
        Integer :: mp, m, l, r, imi, f, total
        Integer*8 :: nb
        Call alloc_spec(s, nf)
        Call alloc_spec(co, nf)
        Allocate(t1(lmax+1))
        total = 0
        Do mp = 1, lmax+1
            m = mvals(mp)
            Allocate(t1(mp)%data(m:lmax, 1:nr, 1:2))
            Do f = 1, nf
                Do imi = 1, 2
                    Do r = 1, nr
                        Do l = m, lmax
                            s(mp)%data(l, r, imi, f) = sin(0.41d0*l + 0.9d0*r + 0.5d0*imi + 1.7d0*f + 0.019d0*m)
                            co(mp)%data(l, r, imi, f) = cos(0.23d0*l + 1.3d0*r + 0.7d0*imi + 0.6d0*f + 0.011d0*m)
                        Enddo
                    Enddo
                Enddo
            Enddo
            total = total + size(s(mp)%data)
        Enddo
        Allocate(packed(total), back(total))
        nb = int(total, 8) * 8
        ds = Device_Allocate(nb)
        dco = Device_Allocate(nb)
        Call pack_spec(s, packed)
        Call Device_Upload(ds, c_loc(packed), nb)
        Call pack_spec(co, packed)
        Call Device_Upload(dco, c_loc(packed), nb)
        Call Hybrid_Output_Final_Device(nr, nf, ds, nf, dco, (/pvar, 1, 1, 1, dpdt_cb, 1, 1, 1, 1/))
        Call Device_Download(c_loc(back), dco, nb)
        Call d_by_dtheta(s, pvar, t1)
        Do mp = 1, lmax+1
            co(mp)%data(lmax, :, :, :) = 0.0d0
            co(mp)%data(:, :, :, dpdt_cb) = t1(mp)%data
        Enddo
        Call pack_spec(co, packed)
        Call Device_Free(ds)
        Call Device_Free(dco)
        Call report('Hybrid_Output_Final_Device (hydro)', maxval(abs(packed - back)), nfail)
    End Subroutine output_final_hydro

    Subroutine pack_spec(s, packed)
        Type(rmcontainer4d), Intent(In) :: s(:)
        Real*8, Intent(Out) :: packed(:)
        Integer :: mp, n, off
        off = 0
        Do mp = 1, size(s)
            n = size(s(mp)%data)
            packed(off + 1:off + n) = Reshape(s(mp)%data, (/ n /))
            off = off + n
        Enddo
    End Subroutine pack_spec

End Program test_sht_raysht
