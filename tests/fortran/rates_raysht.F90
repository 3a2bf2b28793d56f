!
! Rates of the two Fortran call forms of the transform round trip at l_max = 255
! (N_theta = 384, N_phi = 768, N_r = 128, 8 fields, balanced m order):
!   host arrays   Legendre_Transform(s, p), Reform_2a3a, FFT_To_Physical, FFT_To_Spectral,
!                 Reform_3b2b, Legendre_Transform(p, s) on Rayleigh's own host arrays (each
!                 call stages through device memory: the drop-in as it stands)
!   device        the same six steps through the *_Device procedures on buffers that stay in
!                 device memory (inputs uploaded once, before the timed loop)
! Prints one JSON line: milliseconds per round trip of each form (median of the timed
! repetitions after one warm-up) and the host-path result's agreement with the device path.
!
Program rates_raysht
    Use ISO_C_BINDING
    Use Structures
    Use RaySHT_Math_Layer
    Implicit None
    Integer, Parameter :: lmax = 255, nt = 384, nphi = 768, nr = 128, nf = 8, reps = 3
    Integer :: mvals(lmax+1), i, mp, m, off, nl, it
    Type(rmcontainer4d), Allocatable :: s(:)
    Real*8, Allocatable, Target :: p(:,:,:), x(:,:,:,:), packed(:), back(:)
    Type(C_PTR) :: ds, dp, dx
    Integer*8 :: nbs, nbp, nbx, t0, t1, rate
    Real*8 :: host_ms(reps), dev_ms(reps), d

    Do i = 0, (lmax+1)/2 - 1
        mvals(2*i+1) = i
        mvals(2*i+2) = lmax - i
    Enddo
    Call Initialize_Legendre(nt, lmax, mvals, .true.)
    Allocate(s(lmax+1))
    nbs = 0
    Do mp = 1, lmax+1
        m = mvals(mp)
        Allocate(s(mp)%data(m:lmax, 1:nr, 1:2, 1:nf))
        Call fill(s(mp)%data, m)
        nbs = nbs + size(s(mp)%data)
    Enddo
    Allocate(p(nt, 2*nr*nf, lmax+1), x(nphi+2, nr, nt, nf), packed(nbs), back(nbs))
    nbs = nbs * 8
    nbp = int(size(p), 8) * 8
    nbx = int(size(x), 8) * 8
    Call system_clock(count_rate=rate)

    ! host arrays (warm-up, then the timed repetitions)
    Do it = 0, reps
        Call system_clock(t0)
        Call Legendre_Transform(s, p)
        Call Reform_2a3a(p, x)
        Call FFT_To_Physical(x)
        Call FFT_To_Spectral(x)
        Call Reform_3b2b(x, p)
        Call Legendre_Transform(p, s)
        Call system_clock(t1)
        If (it > 0) host_ms(it) = 1.0d3 * (t1 - t0) / rate
        ! the round trip is the identity up to N_phi: restore the input scale
        Do mp = 1, lmax+1
            s(mp)%data = s(mp)%data * nphi
        Enddo
    Enddo

    ! device buffers
    off = 0
    Do mp = 1, lmax+1
        nl = size(s(mp)%data)
        packed(off+1:off+nl) = reshape(s(mp)%data, (/ nl /))
        off = off + nl
    Enddo
    ds = Device_Allocate(nbs)
    dp = Device_Allocate(nbp)
    dx = Device_Allocate(nbx)
    Call Device_Upload(ds, c_loc(packed), nbs)
    Do it = 0, reps
        Call Device_Synchronize()
        Call system_clock(t0)
        Call StP_Device(nr, nf, ds, dp)
        Call Reform_2a3a_Device(nr, nf, nphi, dp, dx)
        Call FFT_To_Physical_Device(nphi, int(nr, 8)*nt*nf, dx)
        Call FFT_To_Spectral_Device(nphi, int(nr, 8)*nt*nf, dx)
        Call Reform_3b2b_Device(nr, nf, nphi, dx, dp)
        Call PtS_Device(nr, nf, dp, ds)
        Call Device_Synchronize()
        Call system_clock(t1)
        If (it > 0) dev_ms(it) = 1.0d3 * (t1 - t0) / rate
        If (it == 0) Then
            ! one host-path round trip on the same input, against the device's
            Call Device_Download(c_loc(back), ds, nbs)
        Endif
    Enddo
    ! the host path's round trip of the same input (packed) equals the device's (back) / N_phi
    Call Legendre_Transform(s, p)
    Call Reform_2a3a(p, x)
    Call FFT_To_Physical(x)
    Call FFT_To_Spectral(x)
    Call Reform_3b2b(x, p)
    Call Legendre_Transform(p, s)
    d = 0.0d0
    off = 0
    Do mp = 1, lmax+1
        nl = size(s(mp)%data)
        d = max(d, maxval(abs(back(off+1:off+nl) - reshape(s(mp)%data, (/ nl /)))))
        off = off + nl
    Enddo
    Call Device_Free(ds)
    Call Device_Free(dp)
    Call Device_Free(dx)
    Call Finalize_Legendre()
    Write(6, '(A, F10.3, A, F10.3, A, ES10.3, A, I0, A, I0, A)') '{"host_arrays_ms": ', median(host_ms), &
        ', "device_buffers_ms": ', median(dev_ms), ', "host_vs_device_max_abs_diff": ', d, &
        ', "host_bytes_staged_per_round_trip": ', 2*(nbs + 2*nbp + 2*nbx), ', "reps": ', reps, '}'

Contains

    Subroutine fill(a, m)
        Real*8, Intent(Out) :: a(m:, :, :, :)
        Integer, Intent(In) :: m
        Integer :: l, r, f
        Do f = 1, nf
            Do r = 1, nr
                Do l = m, lmax
                    a(l, r, 1, f) = sin(0.37d0*l + 1.3d0*r + 0.11d0*m + f) / (l + 1)**2
                    a(l, r, 2, f) = cos(0.23d0*l - 0.7d0*r + 0.19d0*m + f) / (l + 1)**2
                Enddo
            Enddo
        Enddo
        If (m == 0) a(:, :, 2, :) = 0.0d0
    End Subroutine fill

    Function median(v) Result(med)
        Real*8, Intent(In) :: v(:)
        Real*8 :: med, w(size(v)), t
        Integer :: i, j
        w = v
        Do i = 2, size(w)
            t = w(i)
            j = i - 1
            Do While (j >= 1)
                If (w(j) <= t) Exit
                w(j+1) = w(j)
                j = j - 1
            Enddo
            w(j+1) = t
        Enddo
        med = w((size(w) + 1) / 2)
    End Function median

End Program rates_raysht
