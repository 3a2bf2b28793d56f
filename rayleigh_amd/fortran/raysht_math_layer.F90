!
! RaySHT_Math_Layer -- Fortran ISO_C_BINDING layer over libraysht (include/raysht.h).
!
! Gives Rayleigh's Math_Layer callers the interfaces they already use, with the
! same argument types, kinds and ranks (TKR), so that swapping the module is the
! whole integration:
!
!   Initialize_Legendre(nt, lmax, mval, parity)      Legendre_Polynomials.F90:123
!   Legendre_Transform(rmcontainer4d(:), Real*8(:,:,:))   StP_4d_dgp2  (Legendre_Transforms.F90:968)
!   Legendre_Transform(Real*8(:,:,:), rmcontainer4d(:))   PtS_4d_dgpv3 (Legendre_Transforms.F90:1077)
!   Legendre_Transform(rmcontainer(:), Real*8(:,:,:))     StP_2d_dgp   (Legendre_Transforms.F90:588)
!   Legendre_Transform(Real*8(:,:,:), rmcontainer(:))     PtS_2d_dgpv2 (Legendre_Transforms.F90:506)
!   FFT_To_Physical(x, plan, rsc) / FFT_To_Spectral(x, plan, rsc) / FFT_To_Spectral_RSC(x, plan)
!                                                     Fourier_Transform.F90:25-35
!   Type(Cheby_Grid): %Init / %to_spectral / %from_spectral / %dealias_buffer / %d_by_dr_cp /
!       %d_by_dr_nobuffer, with the reference's components  Chebyshev_Polynomials.F90:32-59 (one domain)
!   Cheby_To_Spectral(f_in, c_out) / Cheby_From_Spectral(c_in, f_out) after Cheby_Init(n, rmin, rmax, grid)
!                                                     the same on one module-level grid
!   Cheby_d_by_dr_cp(ind, dind, buffer, dorder)      Chebyshev_Polynomials.F90:560-640
!   Reform_2a3a / Reform_3b2b / Reform_2b1b / Reform_1a2a   SphericalBuffer%reform on one rank
!                                                     (Spherical_Buffer.F90:196-932)
!   d_by_dtheta / d_by_sdtheta / d_by_dphi           Spectral_Derivatives.F90 (rmcontainer4d forms)
!
! The data stay in the caller's host arrays: each call hands their addresses to
! the library's host-array entry points, which stage through device memory.  Callers
! that keep their buffers in device memory use the *_Device procedures (device
! pointers, an optional HIP stream; no transfer) and Device_Allocate / Device_Upload /
! Device_Download / Device_Free / Device_Synchronize for the buffers themselves.
! As in the reference, sizes come from the shapes of the arguments; a device plan
! is built (once) for each (n_r, n_fields) or column count the transforms meet.
! Errors stop the program with the library's message (the reference has no
! error returns either).
!
Module RaySHT_Math_Layer
    Use ISO_C_BINDING
    Use Structures, Only : rmcontainer, rmcontainer3d, rmcontainer4d
    Implicit None
    Private

    Public :: Initialize_Legendre, Finalize_Legendre, Legendre_Transform, Set_Legendre_Precision
    Public :: FFT_To_Physical, FFT_To_Spectral, FFT_To_Spectral_RSC
    Public :: Cheby_Init, Cheby_To_Spectral, Cheby_From_Spectral, Cheby_Finalize, Cheby_d_by_dr_cp
    Public :: Cheby_Grid
    Public :: Reform_2a3a, Reform_3b2b, Reform_2b1b, Reform_1a2a, LM_Count
    Public :: StP_Device, PtS_Device, FFT_To_Physical_Device, FFT_To_Spectral_Device
    Public :: Reform_2a3a_Device, Reform_3b2b_Device, Reform_2b1b_Device, Reform_1a2a_Device
    Public :: Device_Allocate, Device_Free, Device_Upload, Device_Download, Device_Synchronize
    Public :: d_by_dtheta, d_by_sdtheta, d_by_dphi
    Public :: LU_Decompose_full, LU_Solve_full, LU_Decompose_band, LU_Solve_Band
    Public :: Rlm_Spacea_Hydro_Device, Compute_BandCurlB_Device, Adjust_Emf_Device, Rlm_Spaceb_Hydro_Device
    Public :: Hybrid_Output_Initial_Device, Hybrid_Output_Final_Device
    Public :: LM_Scale_Device, Zero_Lmax_Device
    Public :: coloc, gl_weights
    ! arithmetic of the Legendre contractions (enum raysht_precision of include/raysht.h)
    Integer, Parameter, Public :: raysht_fp64 = 0, raysht_bf16x3 = 1, raysht_bf16x6 = 2, &
                                  raysht_int8x3 = 3, raysht_int8x6 = 4

    ! the Gauss-Legendre grid, as the reference exposes it (Real*16, Legendre_Polynomials.F90:28)
    Real*16, Allocatable, Save, Target :: coloc(:), gl_weights(:)

    ! Chebyshev_Polynomials.F90:32-59 for one domain: the same components, the tables kept
    ! on the device behind `handle`
    Type :: Cheby_Grid
        Integer :: domain_count = 0
        Integer :: max_npoly = 0
        Integer :: ntotal = 0
        Integer, Allocatable :: npoly(:), n_x(:), rda(:)
        Real*8, Allocatable :: domain_bounds(:)
        Real*8, Allocatable :: n_even(:), n_odd(:)
        Real*8, Allocatable :: x(:,:), theta(:,:)
        Real*8, Allocatable :: deriv_scaling(:)
        Type(C_PTR) :: handle = C_NULL_PTR
    Contains
        Procedure :: Init => Initialize_Cheby_Grid
        Procedure :: dealias_buffer => dealias_buffer4d
        Procedure :: to_spectral => To_Spectral4D
        Procedure :: from_spectral => From_Spectral4D
        Procedure :: d_by_dr_cp => Cheby_Deriv_Buffer_4D
        Procedure :: d_by_dr_nobuffer => Cheby_Deriv_OutOf_Buffer
        Procedure :: to_spectral_device => To_Spectral_Device
        Procedure :: from_spectral_device => From_Spectral_Device
        Procedure :: finalize => Finalize_Cheby_Grid
    End Type Cheby_Grid

    Integer, Parameter :: max_plans = 8
    Type(C_PTR), Save :: plans(max_plans) = C_NULL_PTR
    Integer, Save :: plan_key(2, max_plans) = -1       ! (n_r, n_fields), or (-ncols, 0) for 2-D
    Integer, Save :: leg_ntheta = 0, leg_lmax = -1, leg_precision = 0
    Integer(C_INT), Allocatable, Save :: leg_mvals(:)
    Type(C_PTR), Save :: cheb = C_NULL_PTR
    Integer, Save :: cheb_n = 0

    Interface
        Function c_plan_create(plan, lmax, ntheta, m_values, n_m, nr, nfields, precision) &
                 Bind(C, name='raysht_plan_create') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Intent(Out) :: plan
            Integer(C_INT), Value :: lmax, ntheta, n_m, nr, nfields, precision
            Integer(C_INT), Intent(In) :: m_values(*)
            Integer(C_INT) :: st
        End Function
        Function c_plan_create_cols(plan, lmax, ntheta, m_values, n_m, ncols, precision) &
                 Bind(C, name='raysht_plan_create_cols') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Intent(Out) :: plan
            Integer(C_INT), Value :: lmax, ntheta, n_m, ncols, precision
            Integer(C_INT), Intent(In) :: m_values(*)
            Integer(C_INT) :: st
        End Function
        Subroutine c_plan_destroy(plan) Bind(C, name='raysht_plan_destroy')
            Import :: C_PTR
            Type(C_PTR), Value :: plan
        End Subroutine
        Function c_plan_nodes(plan, xh, xl, wh, wl) Bind(C, name='raysht_plan_nodes') Result(st)
            Import :: C_PTR, C_INT, C_DOUBLE
            Type(C_PTR), Value :: plan
            Real(C_DOUBLE), Intent(Out) :: xh(*), xl(*), wh(*), wl(*)
            Integer(C_INT) :: st
        End Function
        Function c_stp_host(plan, s_m, p2) Bind(C, name='raysht_legendre_stp_host') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, p2
            Type(C_PTR), Intent(In) :: s_m(*)
            Integer(C_INT) :: st
        End Function
        Function c_pts_host(plan, p2, s_m) Bind(C, name='raysht_legendre_pts_host') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, p2
            Type(C_PTR), Intent(In) :: s_m(*)
            Integer(C_INT) :: st
        End Function
        Function c_fft_phys(nphi, howmany, x) Bind(C, name='raysht_fft_to_physical_host') Result(st)
            Import :: C_PTR, C_INT, C_LONG
            Integer(C_INT), Value :: nphi
            Integer(C_LONG), Value :: howmany
            Type(C_PTR), Value :: x
            Integer(C_INT) :: st
        End Function
        Function c_fft_spec(nphi, howmany, x, rescale) Bind(C, name='raysht_fft_to_spectral_host') Result(st)
            Import :: C_PTR, C_INT, C_LONG
            Integer(C_INT), Value :: nphi, rescale
            Integer(C_LONG), Value :: howmany
            Type(C_PTR), Value :: x
            Integer(C_INT) :: st
        End Function
        Function c_cheb_create(cb, n, rmin, rmax, grid) Bind(C, name='raysht_cheb_create') Result(st)
            Import :: C_PTR, C_INT, C_DOUBLE
            Type(C_PTR), Intent(Out) :: cb
            Integer(C_INT), Value :: n
            Real(C_DOUBLE), Value :: rmin, rmax
            Real(C_DOUBLE), Intent(Out) :: grid(*)
            Integer(C_INT) :: st
        End Function
        Function c_cheb_create_domains(cb, nd, npoly, bounds, dealias_by, grid) &
                Bind(C, name='raysht_cheb_create_domains') Result(st)
            Import :: C_PTR, C_INT, C_DOUBLE
            Type(C_PTR), Intent(Out) :: cb
            Integer(C_INT), Value :: nd
            Integer(C_INT), Intent(In) :: npoly(*)
            Real(C_DOUBLE), Intent(In) :: bounds(*)
            Type(C_PTR), Value :: dealias_by
            Real(C_DOUBLE), Intent(Out) :: grid(*)
            Integer(C_INT) :: st
        End Function
        Subroutine c_cheb_destroy(cb) Bind(C, name='raysht_cheb_destroy')
            Import :: C_PTR
            Type(C_PTR), Value :: cb
        End Subroutine
        Function c_cheb_to_spec(cb, f, c, ncols) Bind(C, name='raysht_cheb_to_spectral_host') Result(st)
            Import :: C_PTR, C_INT, C_LONG
            Type(C_PTR), Value :: cb, f, c
            Integer(C_LONG), Value :: ncols
            Integer(C_INT) :: st
        End Function
        Function c_cheb_from_spec(cb, c, f, ncols) Bind(C, name='raysht_cheb_from_spectral_host') Result(st)
            Import :: C_PTR, C_INT, C_LONG
            Type(C_PTR), Value :: cb, c, f
            Integer(C_LONG), Value :: ncols
            Integer(C_INT) :: st
        End Function
        Function c_cheb_deriv(cb, c, d, ncols, dorder) Bind(C, name='raysht_cheb_deriv_host') Result(st)
            Import :: C_PTR, C_INT, C_LONG
            Type(C_PTR), Value :: cb, c, d
            Integer(C_LONG), Value :: ncols
            Integer(C_INT), Value :: dorder
            Integer(C_INT) :: st
        End Function
        Function c_ang_host(plan, kind, in_m, nf_in, fin, out_m, nf_out, fout) &
            Bind(C, name='raysht_angular_deriv_host') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan
            Integer(C_INT), Value :: kind, nf_in, fin, nf_out, fout
            Type(C_PTR), Intent(In) :: in_m(*), out_m(*)
            Integer(C_INT) :: st
        End Function
        Function c_lu_dec_host(n, kl, ku, mat, lda, pvt, info) Bind(C, name='raysht_lu_decompose_host') Result(st)
            Import :: C_INT, C_PTR
            Integer(C_INT), Value :: n, kl, ku, lda
            Type(C_PTR), Value :: mat, pvt, info
            Integer(C_INT) :: st
        End Function
        Function c_lu_sol_host(n, kl, ku, mat, lda, pvt, rhs, ldb, nrhs) Bind(C, name='raysht_lu_solve_host') Result(st)
            Import :: C_INT, C_PTR
            Integer(C_INT), Value :: n, kl, ku, lda, ldb, nrhs
            Type(C_PTR), Value :: mat, pvt, rhs
            Integer(C_INT) :: st
        End Function
        Function c_plan_nodes_quad(plan, x16, w16) Bind(C, name='raysht_plan_nodes_quad') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, x16, w16
            Integer(C_INT) :: st
        End Function
        Function c_cheb_grid_data(cb, x, theta, w, ds) Bind(C, name='raysht_cheb_grid_data') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: cb, x, theta, w, ds
            Integer(C_INT) :: st
        End Function
        Function c_cheb_to_spec_dev(cb, f, c, ncols, stream) Bind(C, name='raysht_cheb_to_spectral') Result(st)
            Import :: C_PTR, C_INT, C_LONG
            Type(C_PTR), Value :: cb, f, c, stream
            Integer(C_LONG), Value :: ncols
            Integer(C_INT) :: st
        End Function
        Function c_cheb_from_spec_dev(cb, c, f, ncols, stream) Bind(C, name='raysht_cheb_from_spectral') Result(st)
            Import :: C_PTR, C_INT, C_LONG
            Type(C_PTR), Value :: cb, c, f, stream
            Integer(C_LONG), Value :: ncols
            Integer(C_INT) :: st
        End Function
        Function c_p2a3a_host(plan, nphi, p2, p3) Bind(C, name='raysht_permute_2a3a_host') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, p2, p3
            Integer(C_INT), Value :: nphi
            Integer(C_INT) :: st
        End Function
        Function c_p3b2b_host(plan, nphi, p3, p2) Bind(C, name='raysht_permute_3b2b_host') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, p3, p2
            Integer(C_INT), Value :: nphi
            Integer(C_INT) :: st
        End Function
        Function c_p2b1b_host(plan, s_m, p1) Bind(C, name='raysht_permute_2b1b_host') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, p1
            Type(C_PTR), Intent(In) :: s_m(*)
            Integer(C_INT) :: st
        End Function
        Function c_p1a2a_host(plan, p1, s_m) Bind(C, name='raysht_permute_1a2a_host') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, p1
            Type(C_PTR), Intent(In) :: s_m(*)
            Integer(C_INT) :: st
        End Function
        Function c_lm_count(plan) Bind(C, name='raysht_lm_count') Result(n)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan
            Integer(C_INT) :: n
        End Function
        Function c_spacea_hydro(plan, buf, nf, fields, orr, orsq, drho, tor, dlnrho, dsdr, stream) &
                Bind(C, name='raysht_rlm_spacea_hydro') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, buf, orr, orsq, drho, tor, dlnrho, dsdr, stream
            Integer(C_INT), Value :: nf
            Integer(C_INT), Intent(In) :: fields(10)
            Integer(C_INT) :: st
        End Function
        Function c_bandcurlb(plan, buf, nf, fields, o2, stream) Bind(C, name='raysht_compute_bandcurlb') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, buf, o2, stream
            Integer(C_INT), Value :: nf
            Integer(C_INT), Intent(In) :: fields(6)
            Integer(C_INT) :: st
        End Function
        Function c_adjust_emf(plan, buf, nf, et, ep, stream) Bind(C, name='raysht_adjust_emf') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, buf, stream
            Integer(C_INT), Value :: nf, et, ep
            Integer(C_INT) :: st
        End Function
        Function c_output_initial(plan, s, nf, co, nf_co, tmp, fields, o2, stream) &
                Bind(C, name='raysht_hybrid_output_initial') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, s, co, tmp, o2, stream
            Integer(C_INT), Value :: nf, nf_co
            Integer(C_INT), Intent(In) :: fields(6)
            Integer(C_INT) :: st
        End Function
        Function c_output_final(plan, s, nf, co, nf_co, tmp, fields, orr, o2, tor, stream) &
                Bind(C, name='raysht_hybrid_output_final') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, s, co, tmp, orr, o2, tor, stream
            Integer(C_INT), Value :: nf, nf_co
            Integer(C_INT), Intent(In) :: fields(9)
            Integer(C_INT) :: st
        End Function
        Function c_spaceb_hydro(plan, buf, nf, wv, pv, zv, stream) Bind(C, name='raysht_rlm_spaceb_hydro') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, buf, stream
            Integer(C_INT), Value :: nf, wv, pv, zv
            Integer(C_INT) :: st
        End Function
        Function c_lm_scale(plan, a, nf_in, fin, b, nf_out, fout, lfac, rfac, stream) &
                Bind(C, name='raysht_lm_scale') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, a, b, lfac, rfac, stream
            Integer(C_INT), Value :: nf_in, fin, nf_out, fout
            Integer(C_INT) :: st
        End Function
        Function c_zero_lmax(plan, buf, nf, stream) Bind(C, name='raysht_zero_lmax') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, buf, stream
            Integer(C_INT), Value :: nf
            Integer(C_INT) :: st
        End Function
        Function c_stp_dev(plan, s, p2, stream) Bind(C, name='raysht_legendre_stp') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, s, p2, stream
            Integer(C_INT) :: st
        End Function
        Function c_pts_dev(plan, p2, s, stream) Bind(C, name='raysht_legendre_pts') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, p2, s, stream
            Integer(C_INT) :: st
        End Function
        Function c_fft_phys_dev(nphi, howmany, x, stream) Bind(C, name='raysht_fft_to_physical') Result(st)
            Import :: C_PTR, C_INT, C_LONG
            Integer(C_INT), Value :: nphi
            Integer(C_LONG), Value :: howmany
            Type(C_PTR), Value :: x, stream
            Integer(C_INT) :: st
        End Function
        Function c_fft_spec_dev(nphi, howmany, x, rescale, stream) Bind(C, name='raysht_fft_to_spectral') Result(st)
            Import :: C_PTR, C_INT, C_LONG
            Integer(C_INT), Value :: nphi, rescale
            Integer(C_LONG), Value :: howmany
            Type(C_PTR), Value :: x, stream
            Integer(C_INT) :: st
        End Function
        Function c_p2a3a_dev(plan, nphi, p2, p3, stream) Bind(C, name='raysht_permute_2a3a') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, p2, p3, stream
            Integer(C_INT), Value :: nphi
            Integer(C_INT) :: st
        End Function
        Function c_p3b2b_dev(plan, nphi, p3, p2, stream) Bind(C, name='raysht_permute_3b2b') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, p3, p2, stream
            Integer(C_INT), Value :: nphi
            Integer(C_INT) :: st
        End Function
        Function c_p2b1b_dev(plan, s, p1, stream) Bind(C, name='raysht_permute_2b1b') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, s, p1, stream
            Integer(C_INT) :: st
        End Function
        Function c_p1a2a_dev(plan, p1, s, stream) Bind(C, name='raysht_permute_1a2a') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: plan, p1, s, stream
            Integer(C_INT) :: st
        End Function
        ! the HIP runtime's own allocation and copy calls, for device-resident callers
        Function hip_malloc(ptr, nbytes) Bind(C, name='hipMalloc') Result(st)
            Import :: C_PTR, C_INT, C_SIZE_T
            Type(C_PTR), Intent(Out) :: ptr
            Integer(C_SIZE_T), Value :: nbytes
            Integer(C_INT) :: st
        End Function
        Function hip_free(ptr) Bind(C, name='hipFree') Result(st)
            Import :: C_PTR, C_INT
            Type(C_PTR), Value :: ptr
            Integer(C_INT) :: st
        End Function
        Function hip_memcpy(dst, src, nbytes, kind) Bind(C, name='hipMemcpy') Result(st)
            Import :: C_PTR, C_INT, C_SIZE_T
            Type(C_PTR), Value :: dst, src
            Integer(C_SIZE_T), Value :: nbytes
            Integer(C_INT), Value :: kind
            Integer(C_INT) :: st
        End Function
        Function hip_device_synchronize() Bind(C, name='hipDeviceSynchronize') Result(st)
            Import :: C_INT
            Integer(C_INT) :: st
        End Function
        Function c_strerror(status) Bind(C, name='raysht_strerror') Result(msg)
            Import :: C_PTR, C_INT
            Integer(C_INT), Value :: status
            Type(C_PTR) :: msg
        End Function
    End Interface

    Interface Legendre_Transform
        Module Procedure PtS_2d_raysht, StP_2d_raysht
        Module Procedure PtS_4d_raysht, StP_4d_raysht
    End Interface

    Interface FFT_To_Spectral
        Module Procedure r2c_ip4d_raysht
    End Interface

    ! Spectral_Derivatives.F90's generics on the new (rmcontainer4d) layout: field fin -> field
    ! fout of the same buffer (d_dtheta_buffer4D ...), or field fin -> a rmcontainer3d array
    ! (d_dtheta_4dbuff2arr ...).  They use the m values and l_max of Initialize_Legendre.
    Interface d_by_dtheta
        Module Procedure dtheta_buffer4d_raysht, dtheta_4dbuff2arr_raysht
    End Interface
    Interface d_by_sdtheta
        Module Procedure sdtheta_buffer4d_raysht, sdtheta_4dbuff2arr_raysht
    End Interface
    Interface d_by_dphi
        Module Procedure dphi_buffer4d_raysht, dphi_4dbuff2arr_raysht
    End Interface

    Interface FFT_To_Spectral_RSC
        Module Procedure r2c_ip4d_rsc_raysht
    End Interface

    Interface FFT_To_Physical
        Module Procedure c2r_ip4d_raysht
    End Interface

Contains

    Subroutine check(st, what)
        Integer(C_INT), Intent(In) :: st
        Character(len=*), Intent(In) :: what
        Character(kind=C_CHAR), Pointer :: msg(:)
        Integer :: n
        If (st == 0) Return
        Call c_f_pointer(c_strerror(st), msg, [256])
        n = 0
        Do While (n < 256)
            If (msg(n+1) == C_NULL_CHAR) Exit
            n = n + 1
        Enddo
        Write(6, '(A, A, A, I0, A, 256A)') 'RaySHT: ', what, ' failed (status ', st, '): ', msg(1:n)
        Error Stop 1
    End Subroutine check

    ! Initialize_Legendre (Legendre_Polynomials.F90:123): records the grid and m list;
    ! parity is always on (the non-parity branch is not reproduced).
    Subroutine Initialize_Legendre(nt, lmax, mval, parity_in)
        Integer, Intent(In) :: nt, lmax, mval(:)
        Logical, Intent(In) :: parity_in
        Type(C_PTR) :: probe
        If (.not. parity_in) Then
            Write(6, *) 'RaySHT: only the parity-split transforms are provided'
            Error Stop 1
        Endif
        Call Finalize_Legendre()
        leg_ntheta = nt
        leg_lmax = lmax
        Allocate(leg_mvals(size(mval)))
        leg_mvals(:) = mval(:)
        ! the grid, from a one-column plan (the nodes do not depend on the columns)
        Call check(c_plan_create_cols(probe, leg_lmax, leg_ntheta, leg_mvals, size(leg_mvals), 1, 0), &
                   'Initialize_Legendre')
        Allocate(coloc(nt), gl_weights(nt))
        Call check(c_plan_nodes_quad(probe, c_loc(coloc), c_loc(gl_weights)), 'Initialize_Legendre')
        plans(1) = probe
        plan_key(:, 1) = (/ -1, 0 /)
    End Subroutine Initialize_Legendre

    Subroutine Finalize_Legendre()
        Integer :: i
        Do i = 1, max_plans
            If (c_associated(plans(i))) Call c_plan_destroy(plans(i))
            plans(i) = C_NULL_PTR
            plan_key(:, i) = -1
        Enddo
        If (allocated(leg_mvals)) DeAllocate(leg_mvals)
        If (allocated(coloc)) DeAllocate(coloc)
        If (allocated(gl_weights)) DeAllocate(gl_weights)
    End Subroutine Finalize_Legendre

    ! Selects the arithmetic of the following Legendre transforms (raysht_fp64, the
    ! default, or one of the split forms); plans built before are dropped.  No
    ! counterpart in the reference, which always computes in FP64.
    Subroutine Set_Legendre_Precision(precision)
        Integer, Intent(In) :: precision
        Integer :: i
        If (precision < raysht_fp64 .or. precision > raysht_int8x6) Then
            Write(6, *) 'RaySHT: unknown Legendre precision ', precision
            Error Stop 1
        Endif
        leg_precision = precision
        Do i = 1, max_plans
            If (c_associated(plans(i))) Call c_plan_destroy(plans(i))
            plans(i) = C_NULL_PTR
            plan_key(:, i) = -1
        Enddo
    End Subroutine Set_Legendre_Precision

    ! The device plan for (nr, nf) (nf = 0: a 2-D plan over -nr columns), built on first use.
    Function plan_for(nr, nf) Result(plan)
        Integer, Intent(In) :: nr, nf
        Type(C_PTR) :: plan
        Integer :: i
        If (leg_lmax < 0) Then
            Write(6, *) 'RaySHT: Legendre_Transform called before Initialize_Legendre'
            Error Stop 1
        Endif
        Do i = 1, max_plans
            If (plan_key(1, i) == nr .and. plan_key(2, i) == nf .and. c_associated(plans(i))) Then
                plan = plans(i)
                Return
            Endif
        Enddo
        Do i = 1, max_plans
            If (.not. c_associated(plans(i))) Exit
        Enddo
        If (i > max_plans) Then
            i = max_plans
            Call c_plan_destroy(plans(i))
        Endif
        If (nf == 0) Then
            Call check(c_plan_create_cols(plans(i), leg_lmax, leg_ntheta, leg_mvals, size(leg_mvals), -nr, &
                                          leg_precision), 'plan (2-D)')
        Else
            Call check(c_plan_create(plans(i), leg_lmax, leg_ntheta, leg_mvals, size(leg_mvals), nr, nf, &
                                     leg_precision), 'plan')
        Endif
        plan_key(:, i) = (/ nr, nf /)
        plan = plans(i)
    End Function plan_for

    ! StP_4d_dgp2 (Legendre_Transforms.F90:968-1074): data_in(mp)%data(m:lmax, rmin:rmax, 1:2, 1:nf)
    ! -> data_out(1:ntheta, 1:2*nr*nf, 1:n_m)
    Subroutine StP_4d_raysht(data_in, data_out)
        Type(rmcontainer4d), Intent(In), Target :: data_in(:)
        Real*8, Intent(InOut), Target, Contiguous :: data_out(:,:,:)
        Type(C_PTR) :: ptrs(size(data_in))
        Integer :: mp, nr, nf
        nr = size(data_in(1)%data, 2)
        nf = size(data_in(1)%data, 4)
        If (size(data_out, 2) /= 2*nr*nf .or. size(data_out, 3) /= size(leg_mvals) .or. &
            size(data_out, 1) /= leg_ntheta) Then
            Write(6, *) 'RaySHT: StP argument shapes do not match the Legendre setup'
            Error Stop 1
        Endif
        Do mp = 1, size(data_in)
            ptrs(mp) = c_loc(data_in(mp)%data)
        Enddo
        Call check(c_stp_host(plan_for(nr, nf), ptrs, c_loc(data_out)), 'StP')
    End Subroutine StP_4d_raysht

    ! PtS_4d_dgpv3 (Legendre_Transforms.F90:1077-1175)
    Subroutine PtS_4d_raysht(data_in, data_out)
        Real*8, Intent(In), Target, Contiguous :: data_in(:,:,:)
        Type(rmcontainer4d), Intent(InOut), Target :: data_out(1:)
        Type(C_PTR) :: ptrs(size(data_out))
        Integer :: mp, nr, nf
        nr = size(data_out(1)%data, 2)
        nf = size(data_out(1)%data, 4)
        If (size(data_in, 2) /= 2*nr*nf .or. size(data_in, 3) /= size(leg_mvals) .or. &
            size(data_in, 1) /= leg_ntheta) Then
            Write(6, *) 'RaySHT: PtS argument shapes do not match the Legendre setup'
            Error Stop 1
        Endif
        Do mp = 1, size(data_out)
            ptrs(mp) = c_loc(data_out(mp)%data)
        Enddo
        Call check(c_pts_host(plan_for(nr, nf), c_loc(data_in), ptrs), 'PtS')
    End Subroutine PtS_4d_raysht

    ! StP_2d_dgp (Legendre_Transforms.F90:588-668): data_in(mp)%data(m:lmax, 1:nrhs)
    Subroutine StP_2d_raysht(data_in, data_out)
        Type(rmcontainer), Intent(In), Target :: data_in(:)
        Real*8, Intent(InOut), Target, Contiguous :: data_out(:,:,:)
        Type(C_PTR) :: ptrs(size(data_in))
        Integer :: mp
        Do mp = 1, size(data_in)
            ptrs(mp) = c_loc(data_in(mp)%data)
        Enddo
        Call check(c_stp_host(plan_for(-size(data_out, 2), 0), ptrs, c_loc(data_out)), 'StP (2-D)')
    End Subroutine StP_2d_raysht

    ! PtS_2d_dgpv2 (Legendre_Transforms.F90:506-585)
    Subroutine PtS_2d_raysht(data_in, data_out)
        Real*8, Intent(In), Target, Contiguous :: data_in(:,:,:)
        Type(rmcontainer), Intent(InOut), Target :: data_out(1:)
        Type(C_PTR) :: ptrs(size(data_out))
        Integer :: mp
        Do mp = 1, size(data_out)
            ptrs(mp) = c_loc(data_out(mp)%data)
        Enddo
        Call check(c_pts_host(plan_for(-size(data_in, 2), 0), c_loc(data_in), ptrs), 'PtS (2-D)')
    End Subroutine PtS_2d_raysht

    ! c2r_ip4D_fftw (Fourier_Transform.F90:142-183); plan and rsc are accepted and unused,
    ! as in the reference
    Subroutine c2r_ip4d_raysht(x, plan, rsc)
        Real*8, Intent(InOut), Target, Contiguous :: x(:,:,:,:)
        Integer*8, Intent(In), Optional :: plan
        Logical, Intent(In), Optional :: rsc
        Integer(C_LONG) :: howmany
        howmany = int(size(x, 2), C_LONG) * size(x, 3) * size(x, 4)
        Call check(c_fft_phys(int(size(x, 1) - 2, C_INT), howmany, c_loc(x)), 'FFT_To_Physical')
    End Subroutine c2r_ip4d_raysht

    ! r2c_ip4D_fftw (Fourier_Transform.F90:100-140)
    Subroutine r2c_ip4d_raysht(x, plan, rsc)
        Real*8, Intent(InOut), Target, Contiguous :: x(:,:,:,:)
        Integer*8, Intent(In), Optional :: plan
        Logical, Intent(In), Optional :: rsc
        Integer(C_LONG) :: howmany
        howmany = int(size(x, 2), C_LONG) * size(x, 3) * size(x, 4)
        Call check(c_fft_spec(int(size(x, 1) - 2, C_INT), howmany, c_loc(x), 0), 'FFT_To_Spectral')
    End Subroutine r2c_ip4d_raysht

    ! r2c_ip4D_fftw_rsc (Fourier_Transform.F90:56-98)
    Subroutine r2c_ip4d_rsc_raysht(x, plan)
        Real*8, Intent(InOut), Target, Contiguous :: x(:,:,:,:)
        Integer*8, Intent(In), Optional :: plan
        Integer(C_LONG) :: howmany
        howmany = int(size(x, 2), C_LONG) * size(x, 3) * size(x, 4)
        Call check(c_fft_spec(int(size(x, 1) - 2, C_INT), howmany, c_loc(x), 1), 'FFT_To_Spectral_RSC')
    End Subroutine r2c_ip4d_rsc_raysht

    ! Cheby_Grid%init for one domain (Chebyshev_Polynomials.F90:65-207); grid(1:n) receives r.
    Subroutine Cheby_Init(n, rmin, rmax, grid)
        Integer, Intent(In) :: n
        Real*8, Intent(In) :: rmin, rmax
        Real*8, Intent(Out) :: grid(:)
        Call Cheby_Finalize()
        Call check(c_cheb_create(cheb, n, rmin, rmax, grid), 'Cheby_Init')
        cheb_n = n
    End Subroutine Cheby_Init

    Subroutine Cheby_Finalize()
        If (c_associated(cheb)) Call c_cheb_destroy(cheb)
        cheb = C_NULL_PTR
        cheb_n = 0
    End Subroutine Cheby_Finalize

    ! To_Spectral4D (Chebyshev_Polynomials.F90:375-448)
    Subroutine Cheby_To_Spectral(f_in, c_out)
        Real*8, Intent(In), Target, Contiguous :: f_in(:,:,:,:)
        Real*8, Intent(InOut), Target, Contiguous :: c_out(:,:,:,:)
        Integer(C_LONG) :: ncols
        ncols = int(size(f_in, 2), C_LONG) * size(f_in, 3) * size(f_in, 4)
        If (size(f_in, 1) /= cheb_n) Then
            Write(6, *) 'RaySHT: Cheby_To_Spectral radial size does not match Cheby_Init'
            Error Stop 1
        Endif
        Call check(c_cheb_to_spec(cheb, c_loc(f_in), c_loc(c_out), ncols), 'Cheby_To_Spectral')
    End Subroutine Cheby_To_Spectral

    ! From_Spectral4D (Chebyshev_Polynomials.F90:450-558)
    Subroutine Cheby_From_Spectral(c_in, f_out)
        Real*8, Intent(In), Target, Contiguous :: c_in(:,:,:,:)
        Real*8, Intent(InOut), Target, Contiguous :: f_out(:,:,:,:)
        Integer(C_LONG) :: ncols
        ncols = int(size(c_in, 2), C_LONG) * size(c_in, 3) * size(c_in, 4)
        If (size(c_in, 1) /= cheb_n) Then
            Write(6, *) 'RaySHT: Cheby_From_Spectral radial size does not match Cheby_Init'
            Error Stop 1
        Endif
        Call check(c_cheb_from_spec(cheb, c_loc(c_in), c_loc(f_out), ncols), 'Cheby_From_Spectral')
    End Subroutine Cheby_From_Spectral

    ! Cheby_Deriv_Buffer_4D (d_by_dr_cp, Chebyshev_Polynomials.F90:560-640): buffer(:,:,:,dind)
    ! receives the dorder-th radial derivative's coefficients of buffer(:,:,:,ind); ind /= dind
    ! (the reference's in-place branch is unimplemented there).
    Subroutine Cheby_d_by_dr_cp(ind, dind, buffer, dorder)
        Integer, Intent(In) :: ind, dind, dorder
        Real*8, Intent(InOut), Target, Contiguous :: buffer(:,:,:,:)
        Integer(C_LONG) :: ncols
        ncols = int(size(buffer, 2), C_LONG) * size(buffer, 3)
        If (size(buffer, 1) /= cheb_n) Then
            Write(6, *) 'RaySHT: Cheby_d_by_dr_cp radial size does not match Cheby_Init'
            Error Stop 1
        Endif
        If (ind == dind) Then
            Write(6, *) 'RaySHT: Cheby_d_by_dr_cp needs ind /= dind'
            Error Stop 1
        Endif
        Call check(c_cheb_deriv(cheb, c_loc(buffer(1, 1, 1, ind)), c_loc(buffer(1, 1, 1, dind)), ncols, &
                                int(dorder, C_INT)), 'Cheby_d_by_dr_cp')
    End Subroutine Cheby_d_by_dr_cp

    ! ------------------------------------------------------------------ Cheby_Grid
    ! Initialize_Cheby_Grid (Chebyshev_Polynomials.F90:65-207) for ndomains domains of npoly(i)
    ! zeros-grid points between bounds(i) and bounds(i + 1) (ascending, as the caller passes
    ! them) and the optional dealias_by: grid(1:n) receives the radii (descending, n = the total
    ! of npoly), integration_weights(1:n) the radial integration weights; the components are
    ! filled as the reference fills them (domains outermost first, as Initialize_Cheby_Grid
    ! stores them).  dmax and nthread are accepted (the device kernels need neither).
    Subroutine Initialize_Cheby_Grid(self, grid, integration_weights, ndomains, npoly, bounds, dmax, nthread, &
                                     dealias_by, verbose)
        Class(Cheby_Grid) :: self
        Real*8, Intent(InOut) :: grid(1:), integration_weights(1:)
        Real*8, Intent(In) :: bounds(1:)
        Integer, Intent(In) :: ndomains, npoly(1:)
        Integer, Intent(In), Optional :: nthread, dmax
        Integer, Intent(In), Optional, Target :: dealias_by(1:)
        Logical, Intent(In), Optional :: verbose
        Integer :: n, i, j, off, db
        Integer(C_INT), Allocatable, Target :: np_c(:), db_c(:)
        Real(C_DOUBLE), Allocatable :: bd_c(:)
        Type(C_PTR) :: dbp
        Real*8, Allocatable, Target :: x(:), th(:), ds(:), g(:), w(:)
        Logical :: custom
        If (ndomains < 1) Then
            Write(6, *) 'RaySHT: Cheby_Grid%Init needs at least one domain, got ', ndomains
            Error Stop 1
        Endif
        n = sum(npoly(1:ndomains))
        custom = .false.
        If (present(dealias_by)) custom = size(dealias_by) >= ndomains
        If (present(verbose)) Then
            If (verbose) Write(6, *) 'Ndomains: ', ndomains, ' npoly: ', npoly(1:ndomains), &
                                     ' bounds: ', bounds(1:ndomains + 1)
        Endif
        Call self%finalize()
        Allocate(np_c(ndomains), db_c(ndomains), bd_c(ndomains + 1))
        np_c(:) = int(npoly(1:ndomains), C_INT)
        bd_c(:) = bounds(1:ndomains + 1)
        db_c(:) = 0
        dbp = C_NULL_PTR
        If (custom) Then
            db_c(:) = int(dealias_by(1:ndomains), C_INT)
            dbp = c_loc(db_c)
        Endif
        Allocate(x(n), th(n), ds(n), g(n), w(n))
        Call check(c_cheb_create_domains(self%handle, int(ndomains, C_INT), np_c, bd_c, dbp, g), 'Cheby_Grid%Init')
        Call check(c_cheb_grid_data(self%handle, c_loc(x), c_loc(th), c_loc(w), c_loc(ds)), 'Cheby_Grid%Init')
        grid(:) = 0.0d0
        integration_weights(:) = 0.0d0
        grid(1:n) = g(:)
        integration_weights(1:n) = w(:)
        self%domain_count = ndomains
        self%max_npoly = maxval(npoly(1:ndomains))
        self%ntotal = n
        Allocate(self%npoly(ndomains), self%n_x(ndomains), self%rda(ndomains), self%domain_bounds(ndomains + 1), &
                 self%n_even(ndomains), self%n_odd(ndomains))
        Allocate(self%x(self%max_npoly, ndomains), self%theta(self%max_npoly, ndomains), self%deriv_scaling(n))
        self%x(:, :) = 0.0d0
        self%theta(:, :) = 0.0d0
        off = 0
        Do i = 1, ndomains
            j = ndomains - i + 1
            self%npoly(i) = npoly(j)
            self%rda(i) = (2*npoly(j))/3 + 1
            If (custom) Then
                db = dealias_by(j)
                If (db >= 1 .and. db < npoly(j)) self%rda(i) = npoly(j) - db + 1
            Endif
            self%n_odd(i) = npoly(j)/2
            self%n_even(i) = npoly(j)/2 + mod(npoly(j), 2)
            self%n_x(i) = self%n_even(i)
            self%x(1:npoly(j), i) = x(off + 1:off + npoly(j))
            self%theta(1:npoly(j), i) = th(off + 1:off + npoly(j))
            off = off + npoly(j)
        Enddo
        Do i = 1, ndomains + 1
            self%domain_bounds(i) = bounds(ndomains - i + 2)
        Enddo
        self%deriv_scaling(:) = ds(:)
    End Subroutine Initialize_Cheby_Grid

    Subroutine Finalize_Cheby_Grid(self)
        Class(Cheby_Grid) :: self
        If (c_associated(self%handle)) Call c_cheb_destroy(self%handle)
        self%handle = C_NULL_PTR
        If (allocated(self%npoly)) DeAllocate(self%npoly, self%n_x, self%rda, self%domain_bounds, self%n_even, &
                                               self%n_odd, self%x, self%theta, self%deriv_scaling)
        self%domain_count = 0
    End Subroutine Finalize_Cheby_Grid

    Subroutine require_grid(self, n1, what)
        Class(Cheby_Grid), Intent(In) :: self
        Integer, Intent(In) :: n1
        Character(len=*), Intent(In) :: what
        If (.not. c_associated(self%handle)) Then
            Write(6, *) 'RaySHT: ', what, ' before Cheby_Grid%Init'
            Error Stop 1
        Endif
        If (n1 /= self%ntotal) Then
            Write(6, *) 'RaySHT: ', what, ': radial size ', n1, ' does not match the grid (', self%ntotal, ')'
            Error Stop 1
        Endif
    End Subroutine require_grid

    ! To_Spectral4D (Chebyshev_Polynomials.F90:375-448)
    Subroutine To_Spectral4D(self, f_in, c_out)
        Class(Cheby_Grid) :: self
        Real*8, Intent(In), Target, Contiguous :: f_in(:,:,:,:)
        Real*8, Intent(InOut), Target, Contiguous :: c_out(:,:,:,:)
        Call require_grid(self, size(f_in, 1), 'to_spectral')
        Call check(c_cheb_to_spec(self%handle, c_loc(f_in), c_loc(c_out), int(size(f_in) / size(f_in, 1), C_LONG)), &
                   'Cheby_Grid%to_spectral')
    End Subroutine To_Spectral4D

    ! From_Spectral4D (Chebyshev_Polynomials.F90:450-558)
    Subroutine From_Spectral4D(self, c_in, f_out)
        Class(Cheby_Grid) :: self
        Real*8, Intent(In), Target, Contiguous :: c_in(:,:,:,:)
        Real*8, Intent(InOut), Target, Contiguous :: f_out(:,:,:,:)
        Call require_grid(self, size(c_in, 1), 'from_spectral')
        Call check(c_cheb_from_spec(self%handle, c_loc(c_in), c_loc(f_out), int(size(c_in) / size(c_in, 1), C_LONG)), &
                   'Cheby_Grid%from_spectral')
    End Subroutine From_Spectral4D

    ! dealias_buffer4d (Chebyshev_Polynomials.F90:353-374): coefficients rda..npoly of every
    ! domain of every column set to zero, in the caller's (host) array
    Subroutine dealias_buffer4d(self, buffer)
        Class(Cheby_Grid) :: self
        Real*8, Intent(InOut) :: buffer(1:,1:,1:,1:)
        Integer :: h, off
        Call require_grid(self, size(buffer, 1), 'dealias_buffer')
        off = 0
        Do h = 1, self%domain_count
            buffer(off + self%rda(h):off + self%npoly(h), :, :, :) = 0.0d0
            off = off + self%npoly(h)
        Enddo
    End Subroutine dealias_buffer4d

    ! Cheby_Deriv_Buffer_4D (d_by_dr_cp, :560-640): buffer(:,:,:,dind) receives the dorder-th
    ! radial derivative's coefficients of buffer(:,:,:,ind); ind /= dind (the reference's
    ! in-place branch only rescales)
    Subroutine Cheby_Deriv_Buffer_4D(self, ind, dind, buffer, dorder)
        Class(Cheby_Grid) :: self
        Real*8, Intent(InOut), Target, Contiguous :: buffer(0:,1:,1:,1:)
        Integer, Intent(In) :: ind, dind, dorder
        Call require_grid(self, size(buffer, 1), 'd_by_dr_cp')
        If (ind == dind) Then
            Write(6, *) 'RaySHT: Cheby_Grid%d_by_dr_cp needs ind /= dind'
            Error Stop 1
        Endif
        Call check(c_cheb_deriv(self%handle, c_loc(buffer(0, 1, 1, ind)), c_loc(buffer(0, 1, 1, dind)), &
                                int(size(buffer, 2), C_LONG) * size(buffer, 3), int(dorder, C_INT)), &
                   'Cheby_Grid%d_by_dr_cp')
    End Subroutine Cheby_Deriv_Buffer_4D

    ! Cheby_Deriv_OutOf_Buffer (d_by_dr_nobuffer, :646-730): obuffer receives the dorder-th
    ! radial derivative's coefficients of buffer(:,:,:,ind)
    Subroutine Cheby_Deriv_OutOf_Buffer(self, ind, buffer, obuffer, dorder)
        Class(Cheby_Grid) :: self
        Real*8, Intent(InOut), Target, Contiguous :: buffer(0:,1:,1:,1:)
        Real*8, Intent(InOut), Target, Contiguous :: obuffer(0:,1:,1:)
        Integer, Intent(In) :: ind, dorder
        Call require_grid(self, size(buffer, 1), 'd_by_dr_nobuffer')
        Call check(c_cheb_deriv(self%handle, c_loc(buffer(0, 1, 1, ind)), c_loc(obuffer), &
                                int(size(buffer, 2), C_LONG) * size(buffer, 3), int(dorder, C_INT)), &
                   'Cheby_Grid%d_by_dr_nobuffer')
    End Subroutine Cheby_Deriv_OutOf_Buffer

    ! the device-resident forms: ncols columns of npoly values at device pointers f_dev / c_dev
    Subroutine To_Spectral_Device(self, ncols, f_dev, c_dev, stream)
        Class(Cheby_Grid) :: self
        Integer*8, Intent(In) :: ncols
        Type(C_PTR), Intent(In) :: f_dev, c_dev
        Type(C_PTR), Intent(In), Optional :: stream
        Call check(c_cheb_to_spec_dev(self%handle, f_dev, c_dev, int(ncols, C_LONG), opt_stream(stream)), &
                   'Cheby_Grid%to_spectral_device')
    End Subroutine To_Spectral_Device

    Subroutine From_Spectral_Device(self, ncols, c_dev, f_dev, stream)
        Class(Cheby_Grid) :: self
        Integer*8, Intent(In) :: ncols
        Type(C_PTR), Intent(In) :: c_dev, f_dev
        Type(C_PTR), Intent(In), Optional :: stream
        Call check(c_cheb_from_spec_dev(self%handle, c_dev, f_dev, int(ncols, C_LONG), opt_stream(stream)), &
                   'Cheby_Grid%from_spectral_device')
    End Subroutine From_Spectral_Device

    Function opt_stream(stream) Result(st)
        Type(C_PTR), Intent(In), Optional :: stream
        Type(C_PTR) :: st
        st = C_NULL_PTR
        If (present(stream)) st = stream
    End Function opt_stream

    ! ------------------------------------------------------------------ spectral-space passes
    ! The passes of Sphere_Hybrid_Space.F90 on a device buffer of nf fields in the packed
    ! spectral layout of Initialize_Legendre's m values and l_max (data(m:l_max, 1:nr, 1:2,
    ! 1:nf) per m); field arguments are Rayleigh's 1-based field indices (wvar, dwdr, ...),
    ! radial profiles device arrays of nr values.
    ! Velocity_Components + Velocity_Derivatives + d_by_dtheta(tvar, dtdt) (:80,235-355):
    ! fields = (wvar, dwdr, d2wdr2, zvar, dzdr, tvar, vtheta, vphi, dvtdr, dvpdr); dsdr_over_cp
    ! given when pseudo-incompressible
    Subroutine Rlm_Spacea_Hydro_Device(nr, nf, buf, fields, over_rhor, over_rhorsq, drho_term, two_over_r, dlnrho, &
                                       dsdr_over_cp, stream)
        Integer, Intent(In) :: nr, nf, fields(10)
        Type(C_PTR), Intent(In) :: buf, over_rhor, over_rhorsq, drho_term, two_over_r, dlnrho
        Type(C_PTR), Intent(In), Optional :: dsdr_over_cp, stream
        Integer(C_INT) :: f0(10)
        Type(C_PTR) :: ds
        f0 = int(fields - 1, C_INT)
        ds = C_NULL_PTR
        If (present(dsdr_over_cp)) ds = dsdr_over_cp
        Call check(c_spacea_hydro(plan_for(nr, nf), buf, int(nf, C_INT), f0, over_rhor, over_rhorsq, drho_term, &
                                  two_over_r, dlnrho, ds, opt_stream(stream)), 'Rlm_Spacea_Hydro_Device')
    End Subroutine Rlm_Spacea_Hydro_Device

    ! Compute_BandCurlB (:377-460): fields = (cvar, avar, dcdr, dadr, d2cdr2, curlbr)
    Subroutine Compute_BandCurlB_Device(nr, nf, buf, fields, one_over_rsq, stream)
        Integer, Intent(In) :: nr, nf, fields(6)
        Type(C_PTR), Intent(In) :: buf, one_over_rsq
        Type(C_PTR), Intent(In), Optional :: stream
        Integer(C_INT) :: f0(6)
        f0 = int(fields - 1, C_INT)
        Call check(c_bandcurlb(plan_for(nr, nf), buf, int(nf, C_INT), f0, one_over_rsq, opt_stream(stream)), &
                   'Compute_BandCurlB_Device')
    End Subroutine Compute_BandCurlB_Device

    ! Adjust_Emf (:557-602), before rlm_spaceb's final l_max zeroing
    Subroutine Adjust_Emf_Device(nr, nf, buf, emftheta, emfphi, stream)
        Integer, Intent(In) :: nr, nf, emftheta, emfphi
        Type(C_PTR), Intent(In) :: buf
        Type(C_PTR), Intent(In), Optional :: stream
        Call check(c_adjust_emf(plan_for(nr, nf), buf, int(nf, C_INT), int(emftheta - 1, C_INT), &
                                int(emfphi - 1, C_INT), opt_stream(stream)), 'Adjust_Emf_Device')
    End Subroutine Adjust_Emf_Device

    ! Hybrid_Output_Initial's magnetic part (:525-541) from s (nf fields) into the co-buffer (nf_co
    ! fields) and tmp (2 fields, the routine's ftemp3 / ftemp4, kept for the final pass):
    ! fields = (avar, dcdr, d2cdr2, dadr) of s, (avar_cb, dbrdr_cb) of co
    Subroutine Hybrid_Output_Initial_Device(nr, nf, s, nf_co, co, tmp, fields, one_over_rsq, stream)
        Integer, Intent(In) :: nr, nf, nf_co, fields(6)
        Type(C_PTR), Intent(In) :: s, co, tmp, one_over_rsq
        Type(C_PTR), Intent(In), Optional :: stream
        Integer(C_INT) :: f0(6)
        f0 = int(fields - 1, C_INT)
        Call check(c_output_initial(plan_for(nr, nf), s, int(nf, C_INT), co, int(nf_co, C_INT), tmp, f0, &
                                    one_over_rsq, opt_stream(stream)), 'Hybrid_Output_Initial_Device')
    End Subroutine Hybrid_Output_Initial_Device

    ! Hybrid_Output_Final (:543-552) up to its transform: l_max of every co field zeroed,
    ! Hydro_Output_Derivatives (:224-233) and, with tmp (magnetism), BField_Derivatives (:437-498);
    ! fields = (pvar, br, btheta, bphi) of s, (dpdt_cb, dbtdr_cb, dbpdr_cb, dbrdr_cb, dbrdt_cb) of co;
    ! the profiles (One_Over_R, OneOverRSquared, Two_Over_R) are required with tmp
    Subroutine Hybrid_Output_Final_Device(nr, nf, s, nf_co, co, fields, tmp, one_over_r, one_over_rsq, two_over_r, &
                                          stream)
        Integer, Intent(In) :: nr, nf, nf_co, fields(9)
        Type(C_PTR), Intent(In) :: s, co
        Type(C_PTR), Intent(In), Optional :: tmp, one_over_r, one_over_rsq, two_over_r, stream
        Integer(C_INT) :: f0(9)
        Type(C_PTR) :: t, p1, p2, p3
        f0 = int(fields - 1, C_INT)
        t = C_NULL_PTR
        p1 = C_NULL_PTR
        p2 = C_NULL_PTR
        p3 = C_NULL_PTR
        If (present(tmp)) t = tmp
        If (present(one_over_r)) p1 = one_over_r
        If (present(one_over_rsq)) p2 = one_over_rsq
        If (present(two_over_r)) p3 = two_over_r
        Call check(c_output_final(plan_for(nr, nf), s, int(nf, C_INT), co, int(nf_co, C_INT), t, f0, p1, p2, p3, &
                                  opt_stream(stream)), 'Hybrid_Output_Final_Device')
    End Subroutine Hybrid_Output_Final_Device

    ! rlm_spaceb's hydro part (:127-214), its l = 0 and l_max zeroing included
    Subroutine Rlm_Spaceb_Hydro_Device(nr, nf, buf, wvar, pvar, zvar, stream)
        Integer, Intent(In) :: nr, nf, wvar, pvar, zvar
        Type(C_PTR), Intent(In) :: buf
        Type(C_PTR), Intent(In), Optional :: stream
        Call check(c_spaceb_hydro(plan_for(nr, nf), buf, int(nf, C_INT), int(wvar - 1, C_INT), int(pvar - 1, C_INT), &
                                  int(zvar - 1, C_INT), opt_stream(stream)), 'Rlm_Spaceb_Hydro_Device')
    End Subroutine Rlm_Spaceb_Hydro_Device

    ! out(l, r, imi, fout) = (lfac(l) in(l, r, imi, fin)) rfac(r); lfac (l_max + 1) or rfac
    ! (nr) may be C_NULL_PTR
    Subroutine LM_Scale_Device(nr, nf_in, a, fin, nf_out, b, fout, lfac, rfac, stream)
        Integer, Intent(In) :: nr, nf_in, fin, nf_out, fout
        Type(C_PTR), Intent(In) :: a, b, lfac, rfac
        Type(C_PTR), Intent(In), Optional :: stream
        Call check(c_lm_scale(plan_for(nr, nf_in), a, int(nf_in, C_INT), int(fin - 1, C_INT), b, int(nf_out, C_INT), &
                              int(fout - 1, C_INT), lfac, rfac, opt_stream(stream)), 'LM_Scale_Device')
    End Subroutine LM_Scale_Device

    ! "Zero out l_max mode" of rlm_spacea / rlm_spaceb on every field
    Subroutine Zero_Lmax_Device(nr, nf, buf, stream)
        Integer, Intent(In) :: nr, nf
        Type(C_PTR), Intent(In) :: buf
        Type(C_PTR), Intent(In), Optional :: stream
        Call check(c_zero_lmax(plan_for(nr, nf), buf, int(nf, C_INT), opt_stream(stream)), 'Zero_Lmax_Device')
    End Subroutine Zero_Lmax_Device

    ! ------------------------------------------------------------------ reform (one rank)
    ! Transpose_2a3a (Spherical_Buffer.F90:196-459): p2a(1:ntheta, 1:2 nr nf, 1:n_m) ->
    ! p3a(1:nphi+2, 1:nr, 1:ntheta, 1:nf), zero-filled beyond m = l_max
    Subroutine Reform_2a3a(p2a, p3a)
        Real*8, Intent(In), Target, Contiguous :: p2a(:,:,:)
        Real*8, Intent(InOut), Target, Contiguous :: p3a(:,:,:,:)
        Integer :: nr, nf
        nr = size(p3a, 2)
        nf = size(p3a, 4)
        If (size(p2a, 2) /= 2*nr*nf .or. size(p3a, 3) /= leg_ntheta .or. size(p2a, 1) /= leg_ntheta) Then
            Write(6, *) 'RaySHT: Reform_2a3a argument shapes do not match the Legendre setup'
            Error Stop 1
        Endif
        Call check(c_p2a3a_host(plan_for(nr, nf), int(size(p3a, 1) - 2, C_INT), c_loc(p2a), c_loc(p3a)), &
                   'Reform_2a3a')
    End Subroutine Reform_2a3a

    ! Transpose_3b2b (Spherical_Buffer.F90:461-703): p3b -> p2b, m <= l_max read back
    Subroutine Reform_3b2b(p3b, p2b)
        Real*8, Intent(In), Target, Contiguous :: p3b(:,:,:,:)
        Real*8, Intent(InOut), Target, Contiguous :: p2b(:,:,:)
        Integer :: nr, nf
        nr = size(p3b, 2)
        nf = size(p3b, 4)
        If (size(p2b, 2) /= 2*nr*nf .or. size(p3b, 3) /= leg_ntheta .or. size(p2b, 1) /= leg_ntheta) Then
            Write(6, *) 'RaySHT: Reform_3b2b argument shapes do not match the Legendre setup'
            Error Stop 1
        Endif
        Call check(c_p3b2b_host(plan_for(nr, nf), int(size(p3b, 1) - 2, C_INT), c_loc(p3b), c_loc(p2b)), &
                   'Reform_3b2b')
    End Subroutine Reform_3b2b

    ! Transpose_2b1b (Spherical_Buffer.F90:705-843): s2b(mp)%data(m:lmax, 1:nr, 1:2, 1:nf) ->
    ! p1b(1:nr, 1:2, 1:LM_Count(nr, nf), 1:nf), (l, m) pairs in LM_Load_Balance's order
    Subroutine Reform_2b1b(s2b, p1b)
        Type(rmcontainer4d), Intent(In), Target :: s2b(:)
        Real*8, Intent(InOut), Target, Contiguous :: p1b(:,:,:,:)
        Type(C_PTR) :: ptrs(size(s2b))
        Integer :: mp, nr, nf
        nr = size(s2b(1)%data, 2)
        nf = size(s2b(1)%data, 4)
        If (size(p1b, 1) /= nr .or. size(p1b, 4) /= nf .or. size(p1b, 3) /= LM_Count(nr, nf)) Then
            Write(6, *) 'RaySHT: Reform_2b1b argument shapes do not match'
            Error Stop 1
        Endif
        Do mp = 1, size(s2b)
            ptrs(mp) = c_loc(s2b(mp)%data)
        Enddo
        Call check(c_p2b1b_host(plan_for(nr, nf), ptrs, c_loc(p1b)), 'Reform_2b1b')
    End Subroutine Reform_2b1b

    ! Transpose_1a2a (Spherical_Buffer.F90:845-932): the inverse
    Subroutine Reform_1a2a(p1a, s2a)
        Real*8, Intent(In), Target, Contiguous :: p1a(:,:,:,:)
        Type(rmcontainer4d), Intent(InOut), Target :: s2a(:)
        Type(C_PTR) :: ptrs(size(s2a))
        Integer :: mp, nr, nf
        nr = size(s2a(1)%data, 2)
        nf = size(s2a(1)%data, 4)
        If (size(p1a, 1) /= nr .or. size(p1a, 4) /= nf .or. size(p1a, 3) /= LM_Count(nr, nf)) Then
            Write(6, *) 'RaySHT: Reform_1a2a argument shapes do not match'
            Error Stop 1
        Endif
        Do mp = 1, size(s2a)
            ptrs(mp) = c_loc(s2a(mp)%data)
        Enddo
        Call check(c_p1a2a_host(plan_for(nr, nf), c_loc(p1a), ptrs), 'Reform_1a2a')
    End Subroutine Reform_1a2a

    ! my_num_lm of the radial layout: the number of (l, m) pairs of the Legendre setup
    Function LM_Count(nr, nf) Result(n)
        Integer, Intent(In) :: nr, nf
        Integer :: n
        n = int(c_lm_count(plan_for(nr, nf)))
    End Function LM_Count

    ! ------------------------------------------------------------------ device-resident path
    ! The same transforms on device buffers (Type(C_PTR) from Device_Allocate, the layouts of
    ! include/raysht.h: spectral blocks packed in m_values order), asynchronous on `stream`
    ! (default: the null stream); nothing is copied between host and device.
    Subroutine StP_Device(nr, nf, s_dev, p2_dev, stream)
        Integer, Intent(In) :: nr, nf
        Type(C_PTR), Intent(In) :: s_dev, p2_dev
        Type(C_PTR), Intent(In), Optional :: stream
        Call check(c_stp_dev(plan_for(nr, nf), s_dev, p2_dev, opt_stream(stream)), 'StP_Device')
    End Subroutine StP_Device

    Subroutine PtS_Device(nr, nf, p2_dev, s_dev, stream)
        Integer, Intent(In) :: nr, nf
        Type(C_PTR), Intent(In) :: p2_dev, s_dev
        Type(C_PTR), Intent(In), Optional :: stream
        Call check(c_pts_dev(plan_for(nr, nf), p2_dev, s_dev, opt_stream(stream)), 'PtS_Device')
    End Subroutine PtS_Device

    Subroutine FFT_To_Physical_Device(nphi, howmany, x_dev, stream)
        Integer, Intent(In) :: nphi
        Integer*8, Intent(In) :: howmany
        Type(C_PTR), Intent(In) :: x_dev
        Type(C_PTR), Intent(In), Optional :: stream
        Call check(c_fft_phys_dev(int(nphi, C_INT), int(howmany, C_LONG), x_dev, opt_stream(stream)), &
                   'FFT_To_Physical_Device')
    End Subroutine FFT_To_Physical_Device

    Subroutine FFT_To_Spectral_Device(nphi, howmany, x_dev, rsc, stream)
        Integer, Intent(In) :: nphi
        Integer*8, Intent(In) :: howmany
        Type(C_PTR), Intent(In) :: x_dev
        Logical, Intent(In), Optional :: rsc
        Type(C_PTR), Intent(In), Optional :: stream
        Integer(C_INT) :: r
        r = 0
        If (present(rsc)) Then
            If (rsc) r = 1
        Endif
        Call check(c_fft_spec_dev(int(nphi, C_INT), int(howmany, C_LONG), x_dev, r, opt_stream(stream)), &
                   'FFT_To_Spectral_Device')
    End Subroutine FFT_To_Spectral_Device

    Subroutine Reform_2a3a_Device(nr, nf, nphi, p2_dev, p3_dev, stream)
        Integer, Intent(In) :: nr, nf, nphi
        Type(C_PTR), Intent(In) :: p2_dev, p3_dev
        Type(C_PTR), Intent(In), Optional :: stream
        Call check(c_p2a3a_dev(plan_for(nr, nf), int(nphi, C_INT), p2_dev, p3_dev, opt_stream(stream)), &
                   'Reform_2a3a_Device')
    End Subroutine Reform_2a3a_Device

    Subroutine Reform_3b2b_Device(nr, nf, nphi, p3_dev, p2_dev, stream)
        Integer, Intent(In) :: nr, nf, nphi
        Type(C_PTR), Intent(In) :: p3_dev, p2_dev
        Type(C_PTR), Intent(In), Optional :: stream
        Call check(c_p3b2b_dev(plan_for(nr, nf), int(nphi, C_INT), p3_dev, p2_dev, opt_stream(stream)), &
                   'Reform_3b2b_Device')
    End Subroutine Reform_3b2b_Device

    Subroutine Reform_2b1b_Device(nr, nf, s_dev, p1_dev, stream)
        Integer, Intent(In) :: nr, nf
        Type(C_PTR), Intent(In) :: s_dev, p1_dev
        Type(C_PTR), Intent(In), Optional :: stream
        Call check(c_p2b1b_dev(plan_for(nr, nf), s_dev, p1_dev, opt_stream(stream)), 'Reform_2b1b_Device')
    End Subroutine Reform_2b1b_Device

    Subroutine Reform_1a2a_Device(nr, nf, p1_dev, s_dev, stream)
        Integer, Intent(In) :: nr, nf
        Type(C_PTR), Intent(In) :: p1_dev, s_dev
        Type(C_PTR), Intent(In), Optional :: stream
        Call check(c_p1a2a_dev(plan_for(nr, nf), p1_dev, s_dev, opt_stream(stream)), 'Reform_1a2a_Device')
    End Subroutine Reform_1a2a_Device

    ! device buffers for the *_Device procedures (the HIP runtime's hipMalloc / hipMemcpy /
    ! hipFree); errors stop the program
    Function Device_Allocate(nbytes) Result(ptr)
        Integer*8, Intent(In) :: nbytes
        Type(C_PTR) :: ptr
        If (hip_malloc(ptr, int(nbytes, C_SIZE_T)) /= 0) Then
            Write(6, *) 'RaySHT: Device_Allocate of ', nbytes, ' bytes failed'
            Error Stop 1
        Endif
    End Function Device_Allocate

    Subroutine Device_Free(ptr)
        Type(C_PTR), Intent(InOut) :: ptr
        If (c_associated(ptr)) Then
            If (hip_free(ptr) /= 0) Then
                Write(6, *) 'RaySHT: Device_Free failed'
                Error Stop 1
            Endif
        Endif
        ptr = C_NULL_PTR
    End Subroutine Device_Free

    Subroutine Device_Upload(dev, host, nbytes)
        Type(C_PTR), Intent(In) :: dev, host
        Integer*8, Intent(In) :: nbytes
        If (hip_memcpy(dev, host, int(nbytes, C_SIZE_T), 1_C_INT) /= 0) Then
            Write(6, *) 'RaySHT: Device_Upload failed'
            Error Stop 1
        Endif
    End Subroutine Device_Upload

    Subroutine Device_Download(host, dev, nbytes)
        Type(C_PTR), Intent(In) :: host, dev
        Integer*8, Intent(In) :: nbytes
        If (hip_memcpy(host, dev, int(nbytes, C_SIZE_T), 2_C_INT) /= 0) Then
            Write(6, *) 'RaySHT: Device_Download failed'
            Error Stop 1
        Endif
    End Subroutine Device_Download

    Subroutine Device_Synchronize()
        If (hip_device_synchronize() /= 0) Then
            Write(6, *) 'RaySHT: Device_Synchronize failed'
            Error Stop 1
        Endif
    End Subroutine Device_Synchronize

    ! the buffer forms: field fin -> field fout (1-based) of the same rmcontainer4d array
    Subroutine angular_buffer4d(kind, A, fin, fout)
        Integer, Intent(In) :: kind, fin, fout
        Type(rmcontainer4d), Intent(InOut), Target :: A(:)
        Type(C_PTR) :: ptrs(size(A))
        Integer :: mp, nr, nf
        nr = size(A(1)%data, 2)
        nf = size(A(1)%data, 4)
        Do mp = 1, size(A)
            ptrs(mp) = c_loc(A(mp)%data)
        Enddo
        Call check(c_ang_host(plan_for(nr, nf), int(kind, C_INT), ptrs, int(nf, C_INT), int(fin - 1, C_INT), ptrs, &
                              int(nf, C_INT), int(fout - 1, C_INT)), 'angular derivative')
    End Subroutine angular_buffer4d

    ! the 4dbuff2arr forms: field fin of A -> arr(mp)%data(m:lmax, 1:nr, 1:2)
    Subroutine angular_4dbuff2arr(kind, A, fin, arr)
        Integer, Intent(In) :: kind, fin
        Type(rmcontainer4d), Intent(InOut), Target :: A(:)
        Type(rmcontainer3d), Intent(InOut), Target :: arr(:)
        Type(C_PTR) :: pin(size(A)), pout(size(A))
        Integer :: mp, nr, nf
        nr = size(A(1)%data, 2)
        nf = size(A(1)%data, 4)
        Do mp = 1, size(A)
            pin(mp) = c_loc(A(mp)%data)
            pout(mp) = c_loc(arr(mp)%data)
        Enddo
        Call check(c_ang_host(plan_for(nr, nf), int(kind, C_INT), pin, int(nf, C_INT), int(fin - 1, C_INT), pout, &
                              1_C_INT, 0_C_INT), 'angular derivative')
    End Subroutine angular_4dbuff2arr

    ! d_dtheta_buffer4D (Spectral_Derivatives.F90:224-249) / d_dtheta_4dbuff2arr (:329-350)
    Subroutine dtheta_buffer4d_raysht(A, fin, fout)
        Type(rmcontainer4d), Intent(InOut), Target :: A(:)
        Integer, Intent(In) :: fin, fout
        Call angular_buffer4d(0, A, fin, fout)
    End Subroutine dtheta_buffer4d_raysht
    Subroutine dtheta_4dbuff2arr_raysht(A, fin, arr)
        Type(rmcontainer4d), Intent(InOut), Target :: A(:)
        Integer, Intent(In) :: fin
        Type(rmcontainer3d), Intent(InOut), Target :: arr(:)
        Call angular_4dbuff2arr(0, A, fin, arr)
    End Subroutine dtheta_4dbuff2arr_raysht
    ! d_sdtheta_buffer4d (:277-298) / d_sdtheta_4dbuff2arr (:381-401)
    Subroutine sdtheta_buffer4d_raysht(A, fin, fout)
        Type(rmcontainer4d), Intent(InOut), Target :: A(:)
        Integer, Intent(In) :: fin, fout
        Call angular_buffer4d(1, A, fin, fout)
    End Subroutine sdtheta_buffer4d_raysht
    Subroutine sdtheta_4dbuff2arr_raysht(A, fin, arr)
        Type(rmcontainer4d), Intent(InOut), Target :: A(:)
        Integer, Intent(In) :: fin
        Type(rmcontainer3d), Intent(InOut), Target :: arr(:)
        Call angular_4dbuff2arr(1, A, fin, arr)
    End Subroutine sdtheta_4dbuff2arr_raysht
    ! d_by_dphi_rlmbuff4d (:573-603, in place allowed) / d_by_dphi_4dbuff2arr (:495-509)
    Subroutine dphi_buffer4d_raysht(A, fin, fout)
        Type(rmcontainer4d), Intent(InOut), Target :: A(:)
        Integer, Intent(In) :: fin, fout
        Call angular_buffer4d(2, A, fin, fout)
    End Subroutine dphi_buffer4d_raysht
    Subroutine dphi_4dbuff2arr_raysht(A, fin, arr)
        Type(rmcontainer4d), Intent(InOut), Target :: A(:)
        Integer, Intent(In) :: fin
        Type(rmcontainer3d), Intent(InOut), Target :: arr(:)
        Call angular_4dbuff2arr(2, A, fin, arr)
    End Subroutine dphi_4dbuff2arr_raysht

    ! Linear_Solve.F90:1034-1040 LU_Decompose_full (dgetrf) of one matrix, in place
    Subroutine LU_Decompose_full(mat, pvt)
        Real*8, Intent(InOut), Target :: mat(:,:)
        Integer, Intent(Out), Target :: pvt(:)
        Integer(C_INT), Target :: info
        Call check(c_lu_dec_host(Int(Size(mat,2), C_INT), -1_C_INT, -1_C_INT, c_loc(mat), Int(Size(mat,1), C_INT), &
                                 c_loc(pvt), c_loc(info)), 'LU_Decompose_full')
    End Subroutine LU_Decompose_full

    ! Linear_Solve.F90:1087-1107 LU_Solve_full (dgetrs 'N'): na rows (default Size(mat,1)),
    ! nb right-hand sides (default Size(rhs)/Size(rhs,1)), leading dimension Size(rhs,1)
    Subroutine LU_Solve_full(mat, rhs, pvt, na, nb)
        Real*8, Intent(In), Target :: mat(:,:)
        Real*8, Intent(InOut), Target :: rhs(:,:,:)
        Integer, Intent(In), Target :: pvt(:)
        Integer, Optional :: na, nb
        Integer :: ma, mb
        ma = Size(mat,1)
        If (Present(na)) ma = na
        mb = Size(rhs)/Size(rhs,1)
        If (Present(nb)) mb = nb
        Call check(c_lu_sol_host(Int(ma, C_INT), -1_C_INT, -1_C_INT, c_loc(mat), Int(Size(mat,1), C_INT), c_loc(pvt), &
                                 c_loc(rhs), Int(Size(rhs,1), C_INT), Int(mb, C_INT)), 'LU_Solve_full')
    End Subroutine LU_Solve_full

    ! Linear_Solve.F90:1041-1055 LU_Decompose_band (dgbtrf, kl = ku = (ldab - 1) / 3)
    Subroutine LU_Decompose_band(mat, pvt)
        Real*8, Intent(InOut), Target :: mat(:,:)
        Integer, Intent(Out), Target :: pvt(:)
        Integer(C_INT), Target :: info
        Integer(C_INT) :: kl
        kl = Int((Size(mat,1) - 1) / 3, C_INT)
        Call check(c_lu_dec_host(Int(Size(mat,2), C_INT), kl, kl, c_loc(mat), Int(Size(mat,1), C_INT), c_loc(pvt), &
                                 c_loc(info)), 'LU_Decompose_band')
    End Subroutine LU_Decompose_band

    ! Linear_Solve.F90:1056-1086 LU_Solve_Band (dgbtrs 'N'): na rows (default Size(mat,2))
    Subroutine LU_Solve_Band(mat, rhs, pvt, na, nb)
        Real*8, Intent(In), Target :: mat(:,:)
        Real*8, Intent(InOut), Target :: rhs(:,:,:)
        Integer, Intent(In), Target :: pvt(:)
        Integer, Optional :: na, nb
        Integer :: ma, mb
        Integer(C_INT) :: kl
        ma = Size(mat,2)
        If (Present(na)) ma = na
        mb = Size(rhs)/Size(rhs,1)
        If (Present(nb)) mb = nb
        kl = Int((Size(mat,1) - 1) / 3, C_INT)
        Call check(c_lu_sol_host(Int(ma, C_INT), kl, kl, c_loc(mat), Int(Size(mat,1), C_INT), c_loc(pvt), c_loc(rhs), &
                                 Int(Size(rhs,1), C_INT), Int(mb, C_INT)), 'LU_Solve_Band')
    End Subroutine LU_Solve_Band

End Module RaySHT_Math_Layer
