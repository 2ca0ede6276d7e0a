"""ctypes binding of libiwgpu.so (include/iwgpu.h) and a host-side solver handle.

This is the Python face of the C-ABI.  There is no CPU fallback: if the HIP
library is missing or cannot initialise a device, every entry point raises.
"""
import ctypes
import os

import numpy as np

from .params import IWParams

PKG = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(PKG, "libiwgpu.so")

# every symbol declared in include/iwgpu.h
EXPORTS = [
    "iw_params_size", "iw_last_error", "iw_create", "iw_destroy", "iw_set_state", "iw_get_state",
    "iw_get_pressbc", "iw_set_pressbc", "iw_set_dt", "iw_set_coeffs", "iw_get_coeffs", "iw_set_initial", "iw_run", "iw_time",
    "iw_sync", "iw_horfft", "iw_dpdz", "iw_penalty_params", "iw_param_set", "iw_penalty_solve",
    "iw_filter_3dlegenfourier", "iw_interface_avg3d",
    "iw_convec", "iw_calcuvwt", "iw_filtering", "iw_get_sgs",
    "iw_save_restart", "iw_load_restart", "iw_write_restart_file", "iw_read_restart_file", "iw_abvar1", "iw_abvar2", "iw_abvar3", "iw_bdfvarv",
    "iw_bdfvar1", "iw_bdfvar2", "iw_bdfvar3", "iw_set_profiling", "iw_kernel_stats",
    "iw_device_bytes", "iw_maxcalc", "iw_maxcalc_records", "iw_get_dtstate", "iw_set_dtstate",
    "iw_set_meanwave", "iw_get_meanwave", "iw_postp", "iw_speclegen", "iw_set_sgs", "iw_kepe", "iw_kepe_noalloc", "iw_set_mixed", "iw_set_fft_threads", "iw_energy_records",
    "iw_speclegen_records",
    "iw_mlp_nparams", "iw_mlp_create", "iw_mlp_destroy", "iw_mlp_get_params", "iw_mlp_set_params",
    "iw_mlp_forward", "iw_mlp_gradient", "iw_closure_sample", "iw_closure_dataset", "iw_mlp_train",
    "iw_closure_predict", "iw_set_closure",
    "iw_ensemble_create", "iw_ensemble_destroy", "iw_ensemble_run", "iw_ensemble_tune", "iw_ensemble_config",
    "iw_ensemble_set_lanes",
]

MAXCALC_NREC = 42  # IW_MAXCALC_NREC

_lib = None


class IWError(RuntimeError):
    pass


def lib():
    """Load libiwgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise IWError("libiwgpu.so not built: run `python -m iwrepo_amd.build`")
    L = ctypes.CDLL(LIBPATH)
    P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
    L.iw_last_error.restype = ctypes.c_char_p
    L.iw_create.argtypes = [P, ctypes.POINTER(P)]
    L.iw_destroy.argtypes = [P]
    L.iw_set_state.argtypes = [P, I, P]
    L.iw_get_state.argtypes = [P, I, P]
    L.iw_get_pressbc.argtypes = [P, P, P]
    L.iw_set_pressbc.argtypes = [P, P, P]
    L.iw_set_dt.argtypes = [P, D]
    L.iw_set_coeffs.argtypes = [P, D, P, P, D]
    L.iw_get_coeffs.argtypes = [P, P, P, P, P]
    L.iw_set_initial.argtypes = [P, P, P, P, P]
    L.iw_run.argtypes = [P, I, P]
    L.iw_time.restype = D
    L.iw_time.argtypes = [P]
    L.iw_sync.argtypes = [P]
    L.iw_horfft.argtypes = [P, P, I, I]
    L.iw_dpdz.argtypes = [P, P, P]
    L.iw_penalty_params.argtypes = [P, D, I, I, P, P, P, P, P, P]
    L.iw_param_set.argtypes = [P, I, D, D, P, P, P]
    L.iw_penalty_solve.argtypes = [P, D, I, I, D, D, D, D, P, P, P, P]
    L.iw_filter_3dlegenfourier.argtypes = [P, P, I, I]
    L.iw_interface_avg3d.argtypes = [P, P, ctypes.POINTER(I)]
    L.iw_convec.argtypes = [P, I, I]
    L.iw_calcuvwt.argtypes = [P, I]
    L.iw_filtering.argtypes = [P, I]
    L.iw_save_restart.argtypes = [P, ctypes.c_char_p]
    L.iw_load_restart.argtypes = [P, ctypes.c_char_p, D]
    L.iw_write_restart_file.argtypes = [ctypes.c_char_p, P, D, P]
    L.iw_read_restart_file.argtypes = [ctypes.c_char_p, P, P, P, P]
    L.iw_get_sgs.argtypes = [P, I, P]
    L.iw_maxcalc.argtypes = [P, I, P]
    L.iw_maxcalc_records.argtypes = [P, P, I]
    L.iw_get_dtstate.argtypes = [P, P, P, P, P]
    L.iw_set_dtstate.argtypes = [P, D, I, I]
    L.iw_set_meanwave.argtypes = [P, P]
    L.iw_get_meanwave.argtypes = [P, P]
    L.iw_postp.argtypes = [P, P]
    L.iw_set_sgs.argtypes = [P, I, P]
    L.iw_kepe.argtypes = [P, P]
    L.iw_kepe_noalloc.argtypes = [P, P]
    L.iw_set_mixed.argtypes = [P, I]
    L.iw_set_fft_threads.argtypes = [P, I]
    L.iw_speclegen.argtypes = [P, I, P]
    L.iw_energy_records.argtypes = [P, P, I]
    L.iw_speclegen_records.argtypes = [P, P, I]
    for n in ("abvar1", "abvar2", "abvar3", "bdfvarv", "bdfvar1", "bdfvar2", "bdfvar3"):
        f = getattr(L, "iw_" + n)
        f.restype = D
        f.argtypes = [D, D, D]
    L.iw_set_profiling.argtypes = [P, I]
    L.iw_kernel_stats.argtypes = [P, ctypes.c_char_p, I, P, P, P, I]
    L.iw_device_bytes.restype = ctypes.c_longlong
    L.iw_device_bytes.argtypes = [P]
    L.iw_mlp_create.argtypes = [ctypes.c_ulonglong, ctypes.POINTER(P)]
    L.iw_mlp_destroy.argtypes = [P]
    L.iw_mlp_get_params.argtypes = [P, P]
    L.iw_mlp_set_params.argtypes = [P, P]
    L.iw_mlp_forward.argtypes = [P, P, I, P]
    L.iw_mlp_gradient.argtypes = [P, P, P, I, P, P]
    L.iw_closure_sample.argtypes = [P, P, P]
    L.iw_closure_dataset.restype = ctypes.c_longlong
    L.iw_closure_dataset.argtypes = [P, P, P, ctypes.c_longlong]
    L.iw_mlp_train.argtypes = [P, I, I, ctypes.c_float, ctypes.c_ulonglong, P]
    L.iw_closure_predict.argtypes = [P, P, P]
    L.iw_set_closure.argtypes = [P, P]
    L.iw_ensemble_create.argtypes = [P, I, ctypes.POINTER(P)]
    L.iw_ensemble_destroy.argtypes = [P]
    L.iw_ensemble_run.argtypes = [P, I]
    L.iw_ensemble_tune.argtypes = [P, I, P]
    L.iw_ensemble_config.argtypes = [P, P, P]
    L.iw_ensemble_set_lanes.argtypes = [P, I]
    if L.iw_params_size() != ctypes.sizeof(IWParams):
        raise IWError("iw_params layout mismatch")
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise IWError(lib().iw_last_error().decode())


def _ptr(a):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ValueError("arrays must be C-contiguous float64")
    return a.ctypes.data_as(ctypes.c_void_p)


class Solver:
    """One device instance of the step for parameter block ``P``."""

    def __init__(self, P):
        self.L = lib()
        self.P = P
        self.shape = (P.nzloc * P.nsubd, P.ny, P.nx + 2)
        h = ctypes.c_void_p()
        _check(self.L.iw_create(ctypes.byref(P), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.iw_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arr(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != int(np.prod(self.shape)):
            raise ValueError("field has %d values, expected %d" % (a.size, int(np.prod(self.shape))))
        return a

    def set_initial(self, u, v, w, t):
        arrs = [self._arr(a) for a in (u, v, w, t)]
        _check(self.L.iw_set_initial(self.h, *[_ptr(a) for a in arrs]))

    def run(self, nsteps, counts=True):
        c = np.zeros(4 * max(nsteps, 1), dtype=np.int32)
        _check(self.L.iw_run(self.h, nsteps, c.ctypes.data_as(ctypes.c_void_p) if counts else None))
        return c[:4 * nsteps].reshape(nsteps, 4)

    def set_mixed(self, on=True):
        """FP32 physical-space working set in convec (iw_set_mixed)."""
        _check(self.L.iw_set_mixed(self.h, 1 if on else 0))

    def set_fft_threads(self, ft):
        _check(self.L.iw_set_fft_threads(self.h, ft))

    def set_dt(self, dt):
        _check(self.L.iw_set_dt(self.h, ctypes.c_double(dt)))

    def sync(self):
        _check(self.L.iw_sync(self.h))

    def time(self):
        return self.L.iw_time(self.h)

    def get_state(self, idx):
        out = np.empty(self.shape)
        _check(self.L.iw_get_state(self.h, idx, _ptr(out)))
        return out

    def set_state(self, idx, a):
        a = self._arr(a)
        _check(self.L.iw_set_state(self.h, idx, _ptr(a)))

    def state(self):
        return np.stack([self.get_state(i) for i in range(24)])

    def pressbc(self):
        n = self.shape[1] * self.shape[2] * 2
        wnl, wv = np.empty(n), np.empty(n)
        _check(self.L.iw_get_pressbc(self.h, _ptr(wnl), _ptr(wv)))
        return wnl, wv

    def set_coeffs(self, t, fac, bdf, bdfv):
        f = np.ascontiguousarray(fac, dtype=np.float64)
        b = np.ascontiguousarray(bdf, dtype=np.float64)
        _check(self.L.iw_set_coeffs(self.h, t, _ptr(f), _ptr(b), bdfv))

    def horfft(self, b, is_):
        b = np.ascontiguousarray(b, dtype=np.float64).copy()
        nk = b.size // (self.shape[1] * self.shape[2])
        _check(self.L.iw_horfft(self.h, _ptr(b), is_, nk))
        return b

    def dpdz(self, a):
        a = self._arr(a)
        b = np.empty_like(a)
        _check(self.L.iw_dpdz(self.h, _ptr(a), _ptr(b)))
        return b

    def penalty_params(self, eps, ifbcbot, ifbctop):
        """alpha, beta, gamma, delta, tau1, tau2 per subdomain (penalty.f:9-235)."""
        out = [np.zeros(self.P.nsubd) for _ in range(6)]
        _check(self.L.iw_penalty_params(self.h, eps, ifbcbot, ifbctop, *[_ptr(a) for a in out]))
        return out

    def param_set(self, ifbc, eps, ybar):
        """(alpha0, beta0, tau0) of one boundary (penalty.f:57-157)."""
        v = [ctypes.c_double() for _ in range(3)]
        _check(self.L.iw_param_set(self.h, ifbc, eps, ybar, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def penalty_solve(self, eps, ifbcbot, ifbctop, rhsr, rhsi, bc=(0.0, 0.0, 0.0, 0.0)):
        """helmv_penalty on one column pair (rhs already times -eps), on the
        device; bc = (bcrbot, bcibot, bcrtop, bcitop)."""
        rr = np.ascontiguousarray(rhsr, dtype=np.float64)
        ri = np.ascontiguousarray(rhsi, dtype=np.float64)
        fr, fi = np.empty_like(rr), np.empty_like(ri)
        _check(self.L.iw_penalty_solve(self.h, eps, ifbcbot, ifbctop, bc[0], bc[1], bc[2], bc[3], _ptr(rr),
                                       _ptr(ri), _ptr(fr), _ptr(fi)))
        return fr, fi

    def filter3d(self, f, ifourier=1, ilegen=1):
        f = self._arr(f).copy()
        _check(self.L.iw_filter_3dlegenfourier(self.h, _ptr(f), ifourier, ilegen))
        return f

    def interface_avg(self, f):
        f = self._arr(f).copy()
        ic = ctypes.c_int(0)
        _check(self.L.iw_interface_avg3d(self.h, _ptr(f), ctypes.byref(ic)))
        return f, ic.value

    def convec(self, istep, ispecfilternl=1):
        _check(self.L.iw_convec(self.h, istep, ispecfilternl))

    def calcuvwt(self, cpen=1):
        _check(self.L.iw_calcuvwt(self.h, cpen))

    def save_restart(self, path):
        _check(self.L.iw_save_restart(self.h, os.fsencode(path)))

    def load_restart(self, path, cdt=1.0):
        _check(self.L.iw_load_restart(self.h, os.fsencode(path), cdt))

    def filtering(self, isgscalc=0):
        _check(self.L.iw_filtering(self.h, isgscalc))

    def sgs(self, idx):
        out = np.empty(self.shape)
        _check(self.L.iw_get_sgs(self.h, idx, _ptr(out)))
        return out

    def set_sgs(self, idx, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != int(np.prod(self.shape)):
            raise ValueError("SGS tensor: expected nz x ny x nxpp values")
        _check(self.L.iw_set_sgs(self.h, idx, _ptr(a)))

    def maxcalc(self, istep=1):
        """maxcalc (various.f:665-987) on the device state; returns its record."""
        r = np.zeros(MAXCALC_NREC)
        _check(self.L.iw_maxcalc(self.h, istep, _ptr(r)))
        return r

    def maxcalc_records(self, n):
        """(n, MAXCALC_NREC) records of the last run."""
        r = np.zeros((n, MAXCALC_NREC))
        _check(self.L.iw_maxcalc_records(self.h, _ptr(r), n))
        return r

    def postp(self):
        """postp (postp.f:9-76) on the device state: the sixteen values of
        the energy.dat line (tketot, u1tke, u2tke, u3tke, prody, prodz, prod,
        bflux, epsilon, epssgs, epstot, chi, chisgs, chitot, tkemean,
        tempvar); the state takes postp's horfft round trips."""
        r = np.zeros(16)
        _check(self.L.iw_postp(self.h, _ptr(r)))
        return r

    def kepe(self):
        """KE_u, KE_v, KE_w, PE of the device state (CalcE/integration.f:
        14-226 in 3-D; integrcol quadrature); the state is not changed."""
        r = np.zeros(4)
        _check(self.L.iw_kepe(self.h, _ptr(r)))
        return r

    def speclegen(self, idx):
        """speclegencalc (filtering.f:1110-1221) of field idx (0 u, 1 v, 2 w,
        3 temp): (nzloc, nsubd) x-averaged |Legendre coefficients| at
        j = ny/2; the field takes the horfft round trip."""
        r = np.zeros(self.shape[0])
        _check(self.L.iw_speclegen(self.h, idx, _ptr(r)))
        return r.reshape(self.P.nzloc, self.P.nsubd)

    def energy_records(self):
        """(n, 18) postp records (istep, t, the sixteen values) of the last
        iw_run (every npostp steps), or of iw_load_restart."""
        n = self.L.iw_energy_records(self.h, None, 0)
        _check(min(n, 0))
        r = np.zeros((n, 18))
        if n:
            _check(min(self.L.iw_energy_records(self.h, _ptr(r), n), 0))
        return r

    def speclegen_records(self):
        """(n, 2 + 4 nz) speclegencalc records (istep, t, the u, v, w, temp
        spectra, each (nzloc, nsubd) row-major) of the last iw_run."""
        m = 2 + 4 * self.shape[0]
        n = self.L.iw_speclegen_records(self.h, None, 0)
        _check(min(n, 0))
        r = np.zeros((n, m))
        if n:
            _check(min(self.L.iw_speclegen_records(self.h, _ptr(r), n), 0))
        return r

    def set_meanwave(self, mw):
        """/sol2dvar/ arrays, 9 x (nz, nxpp) in COMMON order."""
        a = np.ascontiguousarray(mw, dtype=np.float64)
        if a.size != 9 * self.shape[0] * self.shape[2]:
            raise ValueError("mean-wave arrays: expected 9 x nz x nxpp values")
        _check(self.L.iw_set_meanwave(self.h, _ptr(a)))

    def meanwave(self):
        out = np.zeros((9, self.shape[0], self.shape[2]))
        _check(self.L.iw_get_meanwave(self.h, _ptr(out)))
        return out

    def dtstate(self):
        dt, cdt = ctypes.c_double(), ctypes.c_double()
        iv, idt = ctypes.c_int(), ctypes.c_int()
        _check(self.L.iw_get_dtstate(self.h, ctypes.byref(dt), ctypes.byref(cdt), ctypes.byref(iv),
                                     ctypes.byref(idt)))
        return dt.value, cdt.value, iv.value, idt.value

    def set_profiling(self, on=True):
        _check(self.L.iw_set_profiling(self.h, 1 if on else 0))

    def kernel_stats(self):
        names = ctypes.create_string_buffer(4096)
        avg = np.zeros(64)
        byt = np.zeros(64)
        n = np.zeros(64, dtype=np.int32)
        k = self.L.iw_kernel_stats(self.h, names, 4096, _ptr(avg), n.ctypes.data_as(ctypes.c_void_p),
                                   _ptr(byt), 64)
        if k < 0:
            raise IWError(self.L.iw_last_error().decode())
        nm = names.value.decode().split(";") if k else []
        return {nm[i]: {"avg_ms": float(avg[i]), "launches": int(n[i]), "avg_bytes": float(byt[i])}
                for i in range(k)}

    def device_bytes(self):
        return int(self.L.iw_device_bytes(self.h))


def write_restart_file(path, P, t, state):
    """Host-only writer of the reference restart layout (state: 24 arrays)."""
    st = np.ascontiguousarray(state, dtype=np.float64)
    _check(lib().iw_write_restart_file(os.fsencode(path), ctypes.byref(P), t, _ptr(st)))


def read_restart_file(path, P):
    """Host-only reader: returns (state[24], t, umax)."""
    n = (P.nx + 2) * P.ny * P.nzloc * P.nsubd
    st = np.zeros((24, n))
    t, umax = ctypes.c_double(0.), ctypes.c_double(0.)
    _check(lib().iw_read_restart_file(os.fsencode(path), ctypes.byref(P), ctypes.byref(t), ctypes.byref(umax),
                                      _ptr(st)))
    return st.reshape((24, P.nzloc * P.nsubd, P.ny, P.nx + 2)), t.value, umax.value


def _fptr(a):
    if a.dtype != np.float32 or not a.flags.c_contiguous:
        raise ValueError("arrays must be C-contiguous float32")
    return a.ctypes.data_as(ctypes.c_void_p)


class Closure:
    """The learned sub-grid closure network (iw_mlp_*, iw_closure_*)."""

    NIN, NOUT = 12, 9

    def __init__(self, seed=1):
        self.L = lib()
        h = ctypes.c_void_p()
        _check(self.L.iw_mlp_create(ctypes.c_ulonglong(seed), ctypes.byref(h)))
        self.h = h
        self.nparams = self.L.iw_mlp_nparams()

    def close(self):
        if getattr(self, "h", None):
            self.L.iw_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.nparams, dtype=np.float32)
        _check(self.L.iw_mlp_get_params(self.h, _fptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparams:
            raise ValueError("expected %d values" % self.nparams)
        _check(self.L.iw_mlp_set_params(self.h, _fptr(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((len(x), self.NOUT), dtype=np.float32)
        _check(self.L.iw_mlp_forward(self.h, _fptr(x), len(x), _fptr(y)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.nparams - 2 * (self.NIN + self.NOUT), dtype=np.float32)
        loss = ctypes.c_float()
        _check(self.L.iw_mlp_gradient(self.h, _fptr(x), _fptr(t), len(x), _fptr(g), ctypes.byref(loss)))
        return loss.value, g

    def sample(self, fine, coarse):
        """Coarse-grain fine's state into coarse's and add coarse's samples."""
        _check(self.L.iw_closure_sample(self.h, fine.h, coarse.h))

    def dataset(self, n=None):
        m = self.L.iw_closure_dataset(self.h, None, None, 0)
        if m < 0:
            raise IWError(self.L.iw_last_error().decode())
        n = m if n is None else n
        x = np.zeros((n, self.NIN), dtype=np.float32)
        t = np.zeros((n, self.NOUT), dtype=np.float32)
        if n:
            if self.L.iw_closure_dataset(self.h, _fptr(x), _fptr(t), n) < 0:
                raise IWError(self.L.iw_last_error().decode())
        return x, t

    def train(self, steps, batch=8192, lr=1e-3, seed=7):
        hist = np.zeros(max(steps, 1), dtype=np.float32)
        _check(self.L.iw_mlp_train(self.h, steps, batch, ctypes.c_float(lr), ctypes.c_ulonglong(seed), _fptr(hist)))
        return hist[:steps]

    def predict(self, coarse):
        """tau at coarse's sample points: (nz, ny, nx, 9)."""
        P = coarse.P
        out = np.zeros((P.nzloc * P.nsubd * P.ny * P.nx, self.NOUT), dtype=np.float32)
        _check(self.L.iw_closure_predict(self.h, coarse.h, _fptr(out)))
        return out.reshape(P.nzloc * P.nsubd, P.ny, P.nx, self.NOUT)


def coarse_grain(fine, coarse):
    """Coarse-grain fine's u,v,w,rho into coarse's state (no samples kept)."""
    _check(lib().iw_closure_sample(None, fine.h, coarse.h))


def set_closure(coarse, closure):
    _check(lib().iw_set_closure(coarse.h, closure.h if closure is not None else None))


class Ensemble:
    """Members (Solver objects, initialised) advanced together (iw_ensemble_*)."""

    def __init__(self, members):
        self.L = lib()
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = ctypes.c_void_p()
        _check(self.L.iw_ensemble_create(arr, len(self.members), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.iw_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps):
        _check(self.L.iw_ensemble_run(self.h, nsteps))

    def tune(self, steps=9):
        out = np.zeros(3)
        _check(self.L.iw_ensemble_tune(self.h, steps, out.ctypes.data_as(ctypes.c_void_p)))
        return {"lanes": int(out[0]), "ms_per_step": float(out[1]), "default_ms_per_step": float(out[2])}

    def config(self):
        lanes = ctypes.c_int()
        ft = np.zeros(len(self.members), dtype=np.int32)
        _check(self.L.iw_ensemble_config(self.h, ctypes.byref(lanes), ft.ctypes.data_as(ctypes.c_void_p)))
        return lanes.value, ft

    def set_lanes(self, lanes):
        _check(self.L.iw_ensemble_set_lanes(self.h, lanes))

