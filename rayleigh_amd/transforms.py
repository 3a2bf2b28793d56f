"""Python mirror of Rayleigh's Math_Layer transform interface over the HIP library.

The names follow the reference's Fortran interface so that code (and tests)
written against it read the same:

  LegendrePlan(n_theta, l_max, m_values, n_r, n_fields)   Initialize_Legendre (Legendre_Polynomials.F90:123)
  plan.legendre_transform(s2a, p2a)                        StP_4d_dgp2  (Legendre_Transforms.F90:968)
  plan.legendre_transform(p2b, s2b)                        PtS_4d_dgpv3 (Legendre_Transforms.F90:1077)
  fft_to_physical(x, n_phi) / fft_to_spectral(x, n_phi)    Fourier_Transform.F90:25-35
  fft_to_spectral_rsc(x, n_phi)                            Fourier_Transform.F90:56-98
  plan.reform_2a3a(p2a, p3a) / plan.reform_3b2b(p3b, p2b)  Spherical_Buffer.F90:196-703 (one rank)
  ChebyGrid(n, r_min, r_max).to_spectral / from_spectral / dealias_buffer
                                                           Chebyshev_Polynomials.F90:65-558

All data are float64 tensors on a HIP device in the reference's layouts (see
include/raysht.h); every call is asynchronous on torch's current stream.  There
is no CPU path: the HIP library must be built and loadable.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

PRECISIONS = {"fp64": 0, "bf16x3": 1, "bf16x6": 2, "int8x3": 3, "int8x6": 4}


def _stream(stream=None) -> ctypes.c_void_p:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _dptr(t, name: str, numel: int | None = None) -> ctypes.c_void_p:
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a device tensor")
    if t.dtype != torch.float64:
        raise TypeError(f"{name} must be float64")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name} has {t.numel()} elements, expected {numel}")
    return ctypes.c_void_p(t.data_ptr())


class LegendrePlan:
    """Initialize_Legendre(n_theta, l_max, m_values, parity=.true.) plus the transform
    sizes (n_r radii, n_fields fields) the plan's transforms take."""

    def __init__(self, n_theta: int, l_max: int, m_values, n_r: int, n_fields: int, precision: str = "fp64",
                 n_cols: int | None = None):
        """n_cols instead of (n_r, n_fields) gives the rmcontainer (2-D) pair: blocks
        data(m:l_max, 1:n_cols) (pass n_r = n_fields = 0)."""
        lib = _lib.lib()
        self.n_theta, self.l_max, self.n_r, self.n_fields = int(n_theta), int(l_max), int(n_r), int(n_fields)
        self.m_values = np.ascontiguousarray(m_values, dtype=np.int32)
        self.precision = precision
        h = ctypes.c_void_p()
        mptr = self.m_values.ctypes.data_as(ctypes.c_void_p)
        if n_cols is None:
            self.nrhs = 2 * self.n_r * self.n_fields
            st = lib.raysht_plan_create(ctypes.byref(h), self.l_max, self.n_theta, mptr, len(self.m_values),
                                        self.n_r, self.n_fields, PRECISIONS[precision])
        else:
            self.nrhs = int(n_cols)
            st = lib.raysht_plan_create_cols(ctypes.byref(h), self.l_max, self.n_theta, mptr, len(self.m_values),
                                             self.nrhs, PRECISIONS[precision])
        _lib.check(st, "raysht_plan_create")
        self._h = h
        self.spectral_size = int(lib.raysht_spectral_size(h))
        self.hybrid_size = int(lib.raysht_hybrid_size(h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lib().raysht_plan_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    # --- transforms -------------------------------------------------------
    def stp(self, s, p2, stream=None):
        _lib.check(_lib.lib().raysht_legendre_stp(self._h, _dptr(s, "s", self.spectral_size),
                                                  _dptr(p2, "p2", self.hybrid_size), _stream(stream)),
                   "raysht_legendre_stp")

    def pts(self, p2, s, stream=None):
        _lib.check(_lib.lib().raysht_legendre_pts(self._h, _dptr(p2, "p2", self.hybrid_size),
                                                  _dptr(s, "s", self.spectral_size), _stream(stream)),
                   "raysht_legendre_pts")

    def legendre_transform(self, data_in, data_out, stream=None):
        """The generic of Legendre_Transforms.F90:37-40, resolved by which argument is spectral."""
        if data_in.numel() == self.spectral_size and data_out.numel() == self.hybrid_size:
            return self.stp(data_in, data_out, stream)
        if data_in.numel() == self.hybrid_size and data_out.numel() == self.spectral_size:
            return self.pts(data_in, data_out, stream)
        raise ValueError("legendre_transform: arguments are neither (spectral, hybrid) nor (hybrid, spectral)")

    def reform_2a3a(self, p2a, p3a, n_phi: int, stream=None):
        n = (n_phi + 2) * self.n_r * self.n_theta * self.n_fields
        _lib.check(_lib.lib().raysht_permute_2a3a(self._h, int(n_phi), _dptr(p2a, "p2a", self.hybrid_size),
                                                  _dptr(p3a, "p3a", n), _stream(stream)), "raysht_permute_2a3a")

    def reform_to_physical(self, p2a, p3a, n_phi: int, stream=None):
        """reform_2a3a then fft_to_physical in one call."""
        n = (n_phi + 2) * self.n_r * self.n_theta * self.n_fields
        _lib.check(_lib.lib().raysht_reform_to_physical(self._h, int(n_phi), _dptr(p2a, "p2a", self.hybrid_size),
                                                        _dptr(p3a, "p3a", n), _stream(stream)),
                   "raysht_reform_to_physical")

    def reform_to_spectral(self, p3b, p2b, n_phi: int, stream=None):
        """fft_to_spectral then reform_3b2b in one call (one pass, p3b left unchanged, for
        Rayleigh's lengths N_phi = 3 * 2^k)."""
        n = (n_phi + 2) * self.n_r * self.n_theta * self.n_fields
        _lib.check(_lib.lib().raysht_reform_to_spectral(self._h, int(n_phi), _dptr(p3b, "p3b", n),
                                                        _dptr(p2b, "p2b", self.hybrid_size), _stream(stream)),
                   "raysht_reform_to_spectral")

    def reform_3b2b(self, p3b, p2b, n_phi: int, stream=None):
        n = (n_phi + 2) * self.n_r * self.n_theta * self.n_fields
        _lib.check(_lib.lib().raysht_permute_3b2b(self._h, int(n_phi), _dptr(p3b, "p3b", n),
                                                  _dptr(p2b, "p2b", self.hybrid_size), _stream(stream)),
                   "raysht_permute_3b2b")

    def transpose_2b1b(self, s2b, p1b, stream=None):
        """Transpose_2b1b on one rank: the spectral layout to p1b[f][lm][imi][r], the
        (l, m) pairs in lm_values() order."""
        _lib.check(_lib.lib().raysht_permute_2b1b(self._h, _dptr(s2b, "s2b", self.spectral_size),
                                                  _dptr(p1b, "p1b", self.spectral_size), _stream(stream)),
                   "raysht_permute_2b1b")

    def transpose_1a2a(self, p1a, s2a, stream=None):
        """Transpose_1a2a on one rank: p1a[f][lm][imi][r] back to the spectral layout."""
        _lib.check(_lib.lib().raysht_permute_1a2a(self._h, _dptr(p1a, "p1a", self.spectral_size),
                                                  _dptr(s2a, "s2a", self.spectral_size), _stream(stream)),
                   "raysht_permute_1a2a")

    def angular_derivative(self, kind: str, a, fin: int, out, fout: int, stream=None):
        """Spectral_Derivatives.F90's d_by_dtheta (kind "dtheta"), d_by_sdtheta ("sdtheta") or
        d_by_dphi ("dphi") of field fin of `a` into field fout of `out` (0-based fields; each
        array a whole number of fields of this plan's spectral layout)."""
        if self.n_fields < 1:
            raise ValueError("a plan over anonymous columns has no fields")
        per_field = self.spectral_size // self.n_fields
        if a.numel() % per_field or out.numel() % per_field:
            raise ValueError("arrays must hold whole fields of this plan's spectral layout")
        nf_in, nf_out = a.numel() // per_field, out.numel() // per_field
        if kind == "dphi":
            _lib.check(_lib.lib().raysht_d_by_dphi(self._h, _dptr(a, "a"), nf_in, int(fin), _dptr(out, "out"),
                                                   nf_out, int(fout), _stream(stream)), "raysht_d_by_dphi")
        elif kind in ("dtheta", "sdtheta"):
            _lib.check(_lib.lib().raysht_d_by_dtheta(self._h, 0 if kind == "dtheta" else 1, _dptr(a, "a"), nf_in,
                                                     int(fin), _dptr(out, "out"), nf_out, int(fout),
                                                     _stream(stream)), "raysht_d_by_dtheta")
        else:
            raise ValueError("kind must be dtheta, sdtheta or dphi")

    def _whole_fields(self, *arrays):
        if self.n_fields < 1:
            raise ValueError("a plan over anonymous columns has no fields")
        per_field = self.spectral_size // self.n_fields
        if any(x.numel() % per_field for x in arrays):
            raise ValueError("arrays must hold whole fields of this plan's spectral layout")
        return [x.numel() // per_field for x in arrays]

    def rlm_spaceb_hydro(self, buf, wvar: int, pvar: int, zvar: int, stream=None):
        """Sphere_Hybrid_Space.F90:127-214 rlm_spaceb, hydro part, in place on `buf` (whole
        fields of this plan's spectral layout; W, P, Z the 0-based fields wvar, pvar, zvar):
        W /= l(l+1); Z = (d_by_dphi P - d_by_sdtheta Z) / l(l+1); P = (d_by_sdtheta P +
        d_by_dphi Z) / l(l+1); l = 0 of W, P, Z zeroed for m = 0; l = l_max of every field zeroed."""
        (nf,) = self._whole_fields(buf)
        _lib.check(_lib.lib().raysht_rlm_spaceb_hydro(self._h, _dptr(buf, "buf"), nf, int(wvar), int(pvar), int(zvar),
                                                      _stream(stream)), "raysht_rlm_spaceb_hydro")

    # the radial profiles rlm_spacea_hydro takes, in order
    SPACEA_PROFILES = ("over_rhor", "over_rhorsq", "drho_term", "two_over_r", "dlnrho", "dsdr_over_cp")

    def rlm_spacea_hydro(self, buf, fields, profiles, stream=None):
        """Sphere_Hybrid_Space.F90:235-355 Velocity_Components + Velocity_Derivatives and :80
        d_by_dtheta(tvar, dtdt), in place on `buf` (whole fields).  fields = 0-based (W, dWdr,
        d2Wdr2, Z, dZdr, T, vtheta, vphi, dvtdr, dvpdr); vr, dvrdr, dtdt, dvrdt replace W, dWdr,
        d2Wdr2, dZdr.  profiles: dict of SPACEA_PROFILES float64 device tensors of n_r values
        (dsdr_over_cp None or absent when not pseudo-incompressible)."""
        (nf,) = self._whole_fields(buf)
        v = np.ascontiguousarray(fields, dtype=np.int32)
        if v.size != 10:
            raise ValueError("fields = W, dWdr, d2Wdr2, Z, dZdr, T, vtheta, vphi, dvtdr, dvpdr")
        prof = []
        for k in self.SPACEA_PROFILES:
            x = profiles.get(k)
            if x is None and k != "dsdr_over_cp":
                raise ValueError(f"profile {k} is required")
            prof.append(None if x is None else _dptr(x, k, self.n_r))
        _lib.check(_lib.lib().raysht_rlm_spacea_hydro(self._h, _dptr(buf, "buf"), nf,
                                                      v.ctypes.data_as(ctypes.c_void_p), *prof, _stream(stream)),
                   "raysht_rlm_spacea_hydro")

    def compute_bandcurlb(self, buf, fields, one_over_rsq, stream=None):
        """Sphere_Hybrid_Space.F90:377-460 Compute_BandCurlB, in place on `buf` (whole fields).
        fields = 0-based (C, A, dCdr, dAdr, d2Cdr2, curlbr); Br, Btheta, Bphi, curlbtheta,
        curlbphi replace C, A, dCdr, d2Cdr2, dAdr.  one_over_rsq: float64 device, n_r values."""
        (nf,) = self._whole_fields(buf)
        v = np.ascontiguousarray(fields, dtype=np.int32)
        if v.size != 6:
            raise ValueError("fields = C, A, dCdr, dAdr, d2Cdr2, curlbr")
        _lib.check(_lib.lib().raysht_compute_bandcurlb(self._h, _dptr(buf, "buf"), nf,
                                                       v.ctypes.data_as(ctypes.c_void_p),
                                                       _dptr(one_over_rsq, "one_over_rsq", self.n_r),
                                                       _stream(stream)), "raysht_compute_bandcurlb")

    def adjust_emf(self, buf, emftheta: int, emfphi: int, stream=None):
        """Sphere_Hybrid_Space.F90:557-602 Adjust_Emf, in place on `buf` (whole fields; 0-based
        emftheta, emfphi): emftheta = (d_by_sdtheta emfphi - d_by_dphi emftheta) / l(l+1),
        emfphi = (d_by_dphi emfphi + d_by_sdtheta emftheta) / l(l+1)."""
        (nf,) = self._whole_fields(buf)
        _lib.check(_lib.lib().raysht_adjust_emf(self._h, _dptr(buf, "buf"), nf, int(emftheta), int(emfphi),
                                                _stream(stream)), "raysht_adjust_emf")

    def hybrid_output_initial(self, s, co, tmp, fields, one_over_rsq, stream=None):
        """Sphere_Hybrid_Space.F90:525-541 Hybrid_Output_Initial, magnetic part: from `s` into
        the output co-buffer `co` and the two work fields `tmp` (whole fields each; none
        overlapping).  fields = 0-based (avar, dcdr, d2cdr2, dadr) of s and (avar_cb, dbrdr_cb)
        of co: co[avar_cb] = l(l+1) s[avar] / r^2, co[dbrdr_cb] = s[dcdr], tmp = s[d2cdr2],
        s[dadr].  one_over_rsq: float64 device, n_r values."""
        nf, nf_co, nt = self._whole_fields(s, co, tmp)
        v = np.ascontiguousarray(fields, dtype=np.int32)
        if v.size != 6 or nt != 2:
            raise ValueError("fields = avar, dcdr, d2cdr2, dadr, avar_cb, dbrdr_cb; tmp holds 2 fields")
        _lib.check(_lib.lib().raysht_hybrid_output_initial(
            self._h, _dptr(s, "s"), nf, _dptr(co, "co"), nf_co, _dptr(tmp, "tmp"), v.ctypes.data_as(ctypes.c_void_p),
            _dptr(one_over_rsq, "one_over_rsq", self.n_r), _stream(stream)), "raysht_hybrid_output_initial")

    def hybrid_output_final(self, s, co, fields, tmp=None, one_over_r=None, one_over_rsq=None, two_over_r=None,
                            stream=None):
        """Sphere_Hybrid_Space.F90:543-552 Hybrid_Output_Final up to its transform: l_max of
        every co field zeroed, co[dpdt_cb] = d_by_dtheta s[pvar] and, with `tmp` (the work
        fields of hybrid_output_initial), BField_Derivatives (:437-498) into dbtdr_cb, dbpdr_cb,
        dbrdr_cb (which must hold dcdr) and dbrdt_cb.  fields = 0-based (pvar, br, btheta, bphi)
        of s and (dpdt_cb, dbtdr_cb, dbpdr_cb, dbrdr_cb, dbrdt_cb) of co; the radial profiles
        (float64 device, n_r values) are required with tmp."""
        nf, nf_co = self._whole_fields(s, co)
        v = np.ascontiguousarray(fields, dtype=np.int32)
        if v.size != 9:
            raise ValueError("fields = pvar, br, btheta, bphi, dpdt_cb, dbtdr_cb, dbpdr_cb, dbrdr_cb, dbrdt_cb")
        prof = [None, None, None]
        t = None
        if tmp is not None:
            if self._whole_fields(tmp) != [2]:
                raise ValueError("tmp holds 2 fields")
            if one_over_r is None or one_over_rsq is None or two_over_r is None:
                raise ValueError("one_over_r, one_over_rsq and two_over_r are required with tmp")
            t = _dptr(tmp, "tmp")
            prof = [_dptr(x, k, self.n_r) for x, k in ((one_over_r, "one_over_r"), (one_over_rsq, "one_over_rsq"),
                                                        (two_over_r, "two_over_r"))]
        _lib.check(_lib.lib().raysht_hybrid_output_final(
            self._h, _dptr(s, "s"), nf, _dptr(co, "co"), nf_co, t, v.ctypes.data_as(ctypes.c_void_p), *prof,
            _stream(stream)), "raysht_hybrid_output_final")

    def lm_scale(self, a, fin: int, out, fout: int, lfac=None, rfac=None, stream=None):
        """out(l, r, imi) = (lfac(l) a(l, r, imi)) rfac(r) from field fin of `a` into field fout of
        `out` (the l(l+1) / radial scalings of rlm_spacea and rlm_spaceb); lfac (l_max + 1
        values) and rfac (n_r values) are float64 device tensors, or None for a factor of 1."""
        nf_in, nf_out = self._whole_fields(a, out)
        lf = None if lfac is None else _dptr(lfac, "lfac", self.l_max + 1)
        rf = None if rfac is None else _dptr(rfac, "rfac", self.n_r)
        _lib.check(_lib.lib().raysht_lm_scale(self._h, _dptr(a, "a"), nf_in, int(fin), _dptr(out, "out"), nf_out,
                                              int(fout), lf, rf, _stream(stream)), "raysht_lm_scale")

    def zero_lmax(self, buf, stream=None):
        """data(l_max, :, :, :) = 0 in every m-block of `buf` (whole fields)."""
        (nf,) = self._whole_fields(buf)
        _lib.check(_lib.lib().raysht_zero_lmax(self._h, _dptr(buf, "buf"), nf, _stream(stream)), "raysht_zero_lmax")

    def lm_values(self):
        """LM_Load_Balance's (l, mp) pairs (l_lm_values, mp_lm_values with mp 0-based)."""
        n = _lib.lib().raysht_lm_count(self._h)
        ls, mps = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1), np.int32)
        _lib.check(_lib.lib().raysht_lm_values(self._h, ls.ctypes.data_as(ctypes.c_void_p),
                                               mps.ctypes.data_as(ctypes.c_void_p)), "raysht_lm_values")
        return ls[:n], mps[:n]

    def spectral_to_physical(self, s, work, p3, n_phi: int, stream=None):
        n = (n_phi + 2) * self.n_r * self.n_theta * self.n_fields
        _lib.check(_lib.lib().raysht_spectral_to_physical(self._h, int(n_phi), _dptr(s, "s", self.spectral_size),
                                                          _dptr(work, "work", self.hybrid_size),
                                                          _dptr(p3, "p3", n), _stream(stream)),
                   "raysht_spectral_to_physical")

    def physical_to_spectral(self, p3, work, s, n_phi: int, stream=None):
        n = (n_phi + 2) * self.n_r * self.n_theta * self.n_fields
        _lib.check(_lib.lib().raysht_physical_to_spectral(self._h, int(n_phi), _dptr(p3, "p3", n),
                                                          _dptr(work, "work", self.hybrid_size),
                                                          _dptr(s, "s", self.spectral_size), _stream(stream)),
                   "raysht_physical_to_spectral")

    def round_trip_graph(self, n_phi: int, s_in, work, p3, s_out):
        """The StP -> 2a3a -> c2r -> r2c -> 3b2b -> PtS round trip recorded into a hipGraph
        on these buffers (kept alive by the returned object)."""
        return RoundTripGraph(self, n_phi, s_in, work, p3, s_out)

    # --- setup data, for checking ------------------------------------------
    def nodes(self):
        out = [np.zeros(self.n_theta) for _ in range(4)]
        _lib.check(_lib.lib().raysht_plan_nodes(self._h, *[o.ctypes.data_as(ctypes.c_void_p) for o in out]),
                   "raysht_plan_nodes")
        return tuple(out)

    def table(self, which: int) -> np.ndarray:
        n = int(_lib.lib().raysht_plan_table_size(self._h, which))
        out = np.zeros(max(n, 1))
        _lib.check(_lib.lib().raysht_plan_table(self._h, which, out.ctypes.data_as(ctypes.c_void_p)),
                   "raysht_plan_table")
        return out[:n]

    def tables(self):
        """(p_lm_even, p_lm_odd, ip_lm_even, ip_lm_odd) in the reference's stored orientation."""
        return tuple(self.table(w) for w in range(4))


def _ptr_array(tensors, name: str, sizes):
    return (ctypes.c_void_p * len(tensors))(*[_dptr(t, f"{name}[{i}]", n).value
                                               for i, (t, n) in enumerate(zip(tensors, sizes))])


class LegendreBatch:
    """Up to 64 LegendrePlans of one precision transformed by one grouped launch per
    direction (raysht_batch_*): batch.stp(s_list, p2_list) equals plans[i].stp(s[i],
    p2[i]) for every i."""

    def __init__(self, plans):
        plans = list(plans)
        h = ctypes.c_void_p()
        arr = (ctypes.c_void_p * len(plans))(*[p.handle.value for p in plans])
        _lib.check(_lib.lib().raysht_batch_create(ctypes.byref(h), arr, len(plans)), "raysht_batch_create")
        self._h = h
        self.plans = plans

    def __len__(self):
        return len(self.plans)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lib().raysht_batch_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def _io(self, src, dst, src_sizes, dst_sizes, names):
        if len(src) != len(self.plans) or len(dst) != len(self.plans):
            raise ValueError("one input and one output per plan")
        return _ptr_array(src, names[0], src_sizes), _ptr_array(dst, names[1], dst_sizes)

    def stp(self, s, p2, stream=None):
        a, b = self._io(s, p2, [p.spectral_size for p in self.plans], [p.hybrid_size for p in self.plans],
                        ("s", "p2"))
        _lib.check(_lib.lib().raysht_batch_stp(self._h, a, b, _stream(stream)), "raysht_batch_stp")

    def pts(self, p2, s, stream=None):
        a, b = self._io(p2, s, [p.hybrid_size for p in self.plans], [p.spectral_size for p in self.plans],
                        ("p2", "s"))
        _lib.check(_lib.lib().raysht_batch_pts(self._h, a, b, _stream(stream)), "raysht_batch_pts")

    def round_trip_graph(self, s_in, p2, s_out):
        """The grouped StP -> PtS round trip recorded into a hipGraph."""
        return BatchGraph(self, s_in, p2, s_out)

    def autotune(self, s, p2, s_out, reps: int = 3, stream=None):
        """Times the schedule candidates of both directions on these buffers (outputs are
        overwritten) and keeps the fastest; returns {"stp": {candidate: ms}, "pts": {...}}."""
        ss = [p.spectral_size for p in self.plans]
        hs = [p.hybrid_size for p in self.plans]
        ms = (ctypes.c_double * 8)()
        _lib.check(_lib.lib().raysht_batch_autotune(
            self._h, _ptr_array(s, "s", ss), _ptr_array(p2, "p2", hs), _ptr_array(s_out, "s_out", ss),
            int(reps), _stream(stream), ms), "raysht_batch_autotune")
        return {d: {c: ms[4 * i + c] for c in range(4) if ms[4 * i + c] >= 0.0}
                for i, d in enumerate(("stp", "pts"))}

    def schedule(self):
        """The schedule candidate in use per direction."""
        return {d: _lib.lib().raysht_batch_schedule(self._h, i) for i, d in enumerate(("stp", "pts"))}

    def set_schedule(self, stp: int | None = None, pts: int | None = None):
        for i, c in enumerate((stp, pts)):
            if c is not None:
                _lib.check(_lib.lib().raysht_batch_set_schedule(self._h, i, int(c)), "raysht_batch_set_schedule")


class BatchGraph:
    """raysht_graph_create_batch / raysht_graph_launch."""

    def __init__(self, batch: LegendreBatch, s_in, p2, s_out):
        ss = [p.spectral_size for p in batch.plans]
        hs = [p.hybrid_size for p in batch.plans]
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().raysht_graph_create_batch(
            ctypes.byref(h), batch.handle, _ptr_array(s_in, "s_in", ss), _ptr_array(p2, "p2", hs),
            _ptr_array(s_out, "s_out", ss)), "raysht_graph_create_batch")
        self._h = h
        self._keep = (batch, list(s_in), list(p2), list(s_out))

    def launch(self, stream=None):
        _lib.check(_lib.lib().raysht_graph_launch(self._h, _stream(stream)), "raysht_graph_launch")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lib().raysht_graph_destroy(h)
            self._h = None


class RoundTripGraph:
    """raysht_graph_create_round_trip / raysht_graph_launch."""

    def __init__(self, plan: LegendrePlan, n_phi: int, s_in, work, p3, s_out):
        n3 = (n_phi + 2) * plan.n_r * plan.n_theta * plan.n_fields
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().raysht_graph_create_round_trip(
            ctypes.byref(h), plan.handle, int(n_phi), _dptr(s_in, "s_in", plan.spectral_size),
            _dptr(work, "work", plan.hybrid_size), _dptr(p3, "p3", n3), _dptr(s_out, "s_out", plan.spectral_size)),
            "raysht_graph_create_round_trip")
        self._h = h
        self._keep = (plan, s_in, work, p3, s_out)

    def launch(self, stream=None):
        _lib.check(_lib.lib().raysht_graph_launch(self._h, _stream(stream)), "raysht_graph_launch")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lib().raysht_graph_destroy(h)
            self._h = None


def fft_to_physical(x, n_phi: int, stream=None):
    """c2r_ip4D_fftw: in place on lines of n_phi+2 doubles, unnormalised."""
    if x.numel() % (n_phi + 2):
        raise ValueError("x is not a whole number of lines of n_phi+2")
    _lib.check(_lib.lib().raysht_fft_to_physical(int(n_phi), x.numel() // (n_phi + 2), _dptr(x, "x"),
                                                 _stream(stream)), "raysht_fft_to_physical")


def fft_to_spectral(x, n_phi: int, stream=None, rsc: bool = False):
    """r2c_ip4D_fftw (rsc is accepted and ignored, as in the reference)."""
    del rsc
    if x.numel() % (n_phi + 2):
        raise ValueError("x is not a whole number of lines of n_phi+2")
    _lib.check(_lib.lib().raysht_fft_to_spectral(int(n_phi), x.numel() // (n_phi + 2), _dptr(x, "x"), 0,
                                                 _stream(stream)), "raysht_fft_to_spectral")


def fft_to_spectral_rsc(x, n_phi: int, stream=None):
    """r2c_ip4D_fftw_rsc: r2c, then x(3:) /= n and x(1) /= 2n."""
    if x.numel() % (n_phi + 2):
        raise ValueError("x is not a whole number of lines of n_phi+2")
    _lib.check(_lib.lib().raysht_fft_to_spectral(int(n_phi), x.numel() // (n_phi + 2), _dptr(x, "x"), 1,
                                                 _stream(stream)), "raysht_fft_to_spectral")


def m_values_balanced(l_max: int) -> np.ndarray:
    """The m order of one rank under Load_Balance.F90:126-226 (M_Balance_v0): the pairs
    (i, l_max - i) for i = 0, 1, ..., then the middle m when l_max + 1 is odd."""
    n = l_max + 1
    lo = np.arange(n // 2, dtype=np.int32)
    mv = np.stack([lo, n - 1 - lo], axis=1).ravel()
    return np.concatenate([mv, np.array([n // 2] if n % 2 else [], dtype=np.int32)]).astype(np.int32)


def l_l_plus1(l_max: int) -> np.ndarray:
    """l (l + 1) for l = 0..l_max, from the double l (ProblemSize.F90:506-513)."""
    ell = np.arange(l_max + 1, dtype=np.float64)
    return ell * (ell + 1.0)


def over_l_l_plus1(l_max: int) -> np.ndarray:
    """1 / (l (l + 1)), and 1 at l = 0 (ProblemSize.F90:506-513)."""
    out = np.ones(l_max + 1)
    out[1:] = 1.0 / l_l_plus1(l_max)[1:]
    return out


def _iptr(t, name: str, dtype, numel: int) -> ctypes.c_void_p:
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != dtype or not t.is_contiguous():
        raise TypeError(f"{name} must be a contiguous device tensor of {dtype}")
    if t.numel() < numel:
        raise ValueError(f"{name} has {t.numel()} elements, expected at least {numel}")
    return ctypes.c_void_p(t.data_ptr())


def band_storage(a, kl: int, ku: int | None = None, nlinks: int = 1) -> np.ndarray:
    """Band_Load_Single (Linear_Solve.F90:1358-1409) on the host: an n x n matrix into LAPACK
    band storage ab(2 kl + ku + 1, n) (rows 1..kl left for fill-in, a(r, i) at row
    kl + ku + 1 + r - i of column i; ku = kl as the reference uses).  With nlinks > 1 the
    rows and columns of the nlinks coupled variables (blocks of n / nlinks) are first
    interleaved (block i row r -> row nlinks r + i), which makes a linked system banded."""
    a = np.asarray(a, dtype=np.float64)
    n = a.shape[0]
    ku = kl if ku is None else ku
    if nlinks > 1:
        p = band_interleave(n // nlinks, nlinks)
        a = a[np.ix_(p, p)]
    ab = np.zeros((2 * kl + ku + 1, n), order="F")
    for i in range(n):
        lo, hi = max(0, i - ku), min(n, i + kl + 1)
        ab[kl + ku + np.arange(lo, hi) - i, i] = a[lo:hi, i]
    return ab


def band_interleave(ndim1: int, nlinks: int) -> np.ndarray:
    """The row order of Band_Arrange_RHS (Linear_Solve.F90:1139-1162): new row nlinks r + i takes
    old row r + i ndim1 (0-based), so x_banded = x[band_interleave(...)] and
    x[band_interleave(...)] = x_banded undoes it (Band_ReArrange_RHS)."""
    r, i = np.meshgrid(np.arange(ndim1), np.arange(nlinks), indexing="ij")
    return (r + i * ndim1).ravel()


def lu_decompose_full(mats, pivots, info, n: int, stream=None):
    """LU_Decompose_full (LAPACK dgetrf, Linear_Solve.F90:1034-1045) of every n x n
    column-major matrix in `mats` (float64, nbatch n^2), in place; pivots (int32, nbatch n,
    1-based) and info (int32, nbatch) receive dgetrf's outputs."""
    import torch
    if mats.numel() % (n * n):
        raise ValueError("mats must hold whole n x n matrices")
    nb = mats.numel() // (n * n)
    _lib.check(_lib.lib().raysht_lu_decompose_full(int(n), nb, _dptr(mats, "mats"),
                                                   _iptr(pivots, "pivots", torch.int32, nb * n),
                                                   _iptr(info, "info", torch.int32, nb), _stream(stream)),
               "raysht_lu_decompose_full")


def lu_solve_full(mats, pivots, rhs, rhs_off, nrhs, n: int, max_nrhs: int, stream=None):
    """LU_Solve_full (dgetrs 'N', Linear_Solve.F90:1087-1107): matrix b of the batch solves
    its nrhs[b] columns of n at rhs[rhs_off[b]:] in place (rhs_off int64, nrhs int32)."""
    import torch
    nb = mats.numel() // (n * n)
    _lib.check(_lib.lib().raysht_lu_solve_full(int(n), nb, _dptr(mats, "mats"),
                                               _iptr(pivots, "pivots", torch.int32, nb * n), _dptr(rhs, "rhs"),
                                               _iptr(rhs_off, "rhs_off", torch.int64, nb),
                                               _iptr(nrhs, "nrhs", torch.int32, nb), int(max_nrhs), _stream(stream)),
               "raysht_lu_solve_full")


def lu_decompose_band(mats, pivots, info, n: int, kl: int, ku: int | None = None, ldab: int | None = None,
                      stream=None):
    """LU_Decompose_band (LAPACK dgbtrf, Linear_Solve.F90:1041-1055) of every band matrix in
    `mats` (float64, nbatch ldab n, LAPACK band storage as band_storage builds it; ku = kl and
    ldab = 2 kl + ku + 1 by default), in place; pivots (int32, nbatch n, 1-based), info
    (int32, nbatch)."""
    import torch
    ku = kl if ku is None else ku
    ldab = 2 * kl + ku + 1 if ldab is None else ldab
    if mats.numel() % (ldab * n):
        raise ValueError("mats must hold whole ldab x n band matrices")
    nb = mats.numel() // (ldab * n)
    _lib.check(_lib.lib().raysht_lu_decompose_band(int(n), int(kl), int(ku), nb, _dptr(mats, "mats"), int(ldab),
                                                   _iptr(pivots, "pivots", torch.int32, nb * n),
                                                   _iptr(info, "info", torch.int32, nb), _stream(stream)),
               "raysht_lu_decompose_band")


def lu_solve_band(mats, pivots, rhs, rhs_off, nrhs, n: int, kl: int, max_nrhs: int, ku: int | None = None,
                  ldab: int | None = None, stream=None):
    """LU_Solve_Band (dgbtrs 'N', Linear_Solve.F90:1056-1086): matrix b of the batch solves its
    nrhs[b] columns of n at rhs[rhs_off[b]:] in place."""
    import torch
    ku = kl if ku is None else ku
    ldab = 2 * kl + ku + 1 if ldab is None else ldab
    nb = mats.numel() // (ldab * n)
    _lib.check(_lib.lib().raysht_lu_solve_band(int(n), int(kl), int(ku), nb, _dptr(mats, "mats"), int(ldab),
                                               _iptr(pivots, "pivots", torch.int32, nb * n), _dptr(rhs, "rhs"),
                                               _iptr(rhs_off, "rhs_off", torch.int64, nb),
                                               _iptr(nrhs, "nrhs", torch.int32, nb), int(max_nrhs), _stream(stream)),
               "raysht_lu_solve_band")


class ChebyGrid:
    """Cheby_Grid with one domain of n zeros-grid points on [r_min, r_max], or (ChebyGrid.domains)
    several domains; n is then the total of every domain's points."""

    def __init__(self, n: int, r_min: float, r_max: float):
        self.n = int(n)
        self.grid = np.zeros(self.n)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().raysht_cheb_create(ctypes.byref(h), self.n, float(r_min), float(r_max),
                                                 self.grid.ctypes.data_as(ctypes.c_void_p)), "raysht_cheb_create")
        self._h = h

    @classmethod
    def domains(cls, npoly, bounds, dealias_by=None):
        """Cheby_Grid%init with len(npoly) domains (Chebyshev_Polynomials.F90:65-207): npoly and
        bounds (len(npoly) + 1) in ascending order of radius, as the reference takes them;
        dealias_by optional, same order.  The grid holds the domains outermost first."""
        self = cls.__new__(cls)
        npl = np.ascontiguousarray(npoly, dtype=np.int32)
        bd = np.ascontiguousarray(bounds, dtype=np.float64)
        if bd.size != npl.size + 1:
            raise ValueError("bounds has one more value than npoly")
        db = None if dealias_by is None else np.ascontiguousarray(dealias_by, dtype=np.int32)
        if db is not None and db.size != npl.size:
            raise ValueError("dealias_by has one value per domain")
        self.n = int(npl.sum())
        self.grid = np.zeros(self.n)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().raysht_cheb_create_domains(
            ctypes.byref(h), int(npl.size), npl.ctypes.data_as(ctypes.c_void_p), bd.ctypes.data_as(ctypes.c_void_p),
            None if db is None else db.ctypes.data_as(ctypes.c_void_p), self.grid.ctypes.data_as(ctypes.c_void_p)),
            "raysht_cheb_create_domains")
        self._h = h
        return self

    def grid_data(self):
        """(x, theta, weights, deriv_scaling), n values each"""
        out = [np.zeros(self.n) for _ in range(4)]
        _lib.check(_lib.lib().raysht_cheb_grid_data(self._h, *[o.ctypes.data_as(ctypes.c_void_p) for o in out]),
                   "raysht_cheb_grid_data")
        return tuple(out)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lib().raysht_cheb_destroy(h)
            self._h = None

    FORMS = {"dct": 0, "matrix": 1}

    def set_form(self, form: str):
        """"dct" (power-of-two n, 16 <= n <= 256; the default for n = 128 and 256): DCT-II / DCT-III through
        one complex FFT per column, with the first-order correction for the accumulated
        angles of the reference's tables; "matrix": the parity-split products with the
        Gen_Tn tables (Chebyshev_Polynomials.F90:375-558)."""
        _lib.check(_lib.lib().raysht_cheb_set_form(self._h, self.FORMS[form]), "raysht_cheb_set_form")

    def tables(self):
        nx = self.n - self.n // 2
        ce, co = np.zeros(nx * nx), np.zeros(max(nx * (self.n // 2), 1))
        _lib.check(_lib.lib().raysht_cheb_table(self._h, 0, ce.ctypes.data_as(ctypes.c_void_p)), "cheb_table")
        _lib.check(_lib.lib().raysht_cheb_table(self._h, 1, co.ctypes.data_as(ctypes.c_void_p)), "cheb_table")
        return ce, co[: nx * (self.n // 2)]

    def to_spectral(self, f_in, c_out, stream=None):
        if f_in.numel() % self.n or f_in.numel() != c_out.numel():
            raise ValueError("buffers must hold whole columns of n and match")
        _lib.check(_lib.lib().raysht_cheb_to_spectral(self._h, _dptr(f_in, "f_in"), _dptr(c_out, "c_out"),
                                                      f_in.numel() // self.n, _stream(stream)),
                   "raysht_cheb_to_spectral")

    def from_spectral(self, c_in, f_out, stream=None):
        if c_in.numel() % self.n or c_in.numel() != f_out.numel():
            raise ValueError("buffers must hold whole columns of n and match")
        _lib.check(_lib.lib().raysht_cheb_from_spectral(self._h, _dptr(c_in, "c_in"), _dptr(f_out, "f_out"),
                                                        c_in.numel() // self.n, _stream(stream)),
                   "raysht_cheb_from_spectral")

    def dealias_buffer(self, c, stream=None):
        if c.numel() % self.n:
            raise ValueError("buffer must hold whole columns of n")
        _lib.check(_lib.lib().raysht_cheb_dealias(self._h, _dptr(c, "c"), c.numel() // self.n, _stream(stream)),
                   "raysht_cheb_dealias")

    def d_by_dr(self, c_in, c_out, dorder=1, stream=None):
        """Cheby_Deriv_OutOf_Buffer (d_by_dr_nobuffer, Chebyshev_Polynomials.F90:642-730):
        c_out = coefficients of the dorder-th radial derivative of every column of c_in."""
        if c_in.numel() % self.n or c_in.numel() != c_out.numel():
            raise ValueError("buffers must hold whole columns of n and match")
        _lib.check(_lib.lib().raysht_cheb_deriv(self._h, _dptr(c_in, "c_in"), _dptr(c_out, "c_out"),
                                                c_in.numel() // self.n, int(dorder), _stream(stream)),
                   "raysht_cheb_deriv")

    def d_by_dr_cp(self, ind, dind, buffer, dorder=1, stream=None):
        """Cheby_Deriv_Buffer_4D (d_by_dr_cp, :560-640): buffer is the C-order view of the
        reference's buffer(0:n-1, :, :, 1:nvar), i.e. slowest index = variable; variable
        dind (1-based) receives the dorder-th derivative of variable ind."""
        if not (1 <= ind <= buffer.shape[0] and 1 <= dind <= buffer.shape[0]) or ind == dind:
            raise ValueError("ind and dind must name two different variables of the buffer")
        self.d_by_dr(buffer[ind - 1], buffer[dind - 1], dorder, stream)
