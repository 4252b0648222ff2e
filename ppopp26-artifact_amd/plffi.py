"""ctypes binding of libpl's C-ABI (include/pl.h) for tests and bench.py.

Device memory and streams come from PyTorch (plumbing only): a field vector
is a torch int64 tensor of shape (n, 4) on the GPU, i.e. n canonical
Montgomery elements of 32 bytes.  All compute goes through lib/libpl.so; if
that library is missing the import fails loudly (there is no fallback).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpl.so")

PL_OK = 0
STATUS = {0: "ok", -1: "invalid argument", -2: "HIP runtime error", -3: "unsupported size",
          -4: "device out of memory", -5: "no HIP device"}

_lib = None
_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32
_i = ctypes.c_int


class PlError(RuntimeError):
    def __init__(self, fn, status):
        detail = ""
        if status in (-2, -4) and _lib is not None:
            detail = ": " + _lib.pl_last_hip_error().decode()
        super().__init__("%s failed: %s (%d)%s" % (fn, STATUS.get(status, "?"), status, detail))
        self.status = status


# name -> (restype, argtypes); the symbol list a loader test checks
SIGNATURES = {
    "pl_status_string": (ctypes.c_char_p, [_i]),
    "pl_version": (_i, []),
    "pl_last_hip_error": (ctypes.c_char_p, []),
    "pl_domain_create": (_i, [_i, ctypes.POINTER(_vp)]),
    "pl_domain_destroy": (_i, [_vp]),
    "pl_domain_lg": (_i, [_vp]),
    "pl_params_fill": (_i, [_vp, _vp]),
    "pl_domain_bytes": (ctypes.c_size_t, [_vp]),
    "pl_ntt": (_i, [_vp, _vp, _i, _i, _vp, _vp, _vp]),
    "pl_ntt_batch": (_i, [_vp, _vp, _i, _u32, _i, _vp, _vp, _vp]),
    "pl_ntt_work_bytes": (ctypes.c_size_t, [_i, _u32]),
    "pl_ntt_num_passes": (_i, [_i]),
    "pl_domain_num_passes": (_i, [_vp]),
    "pl_ntt_pass": (_i, [_vp, _vp, _i, _u32, _i, _i, _vp, _vp]),
    "pl_ntt_stage_no_rotate": (_i, [_vp, _u64, _i, _i, _i, _u64, _vp, _vp]),
    "pl_ct_stage_indices": (None, [_u64, _i, _i, ctypes.POINTER(_u64), ctypes.POINTER(_u64)]),
    "pl_rotate_stage_output_index": (_u64, [_u64, _i, _i]),
    "pl_stage_gather": (_i, [_vp, _vp, _u64, _u64, _i, _i, _vp]),
    "pl_stage_scatter": (_i, [_vp, _vp, _u64, _u64, _i, _i, _i, _vp]),
    "pl_stage_transition": (_i, [_vp, _vp, _u64, _u64, _i, _i, _i, _i, _vp]),
    "pl_bit_reverse": (_i, [_vp, _vp, _i, _vp]),
    "pl_ntt_segmented_work_bytes": (ctypes.c_size_t, [_i, _i, _i]),
    "pl_ntt_segmented": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp]),
    "pl_ntt_host_work_bytes": (ctypes.c_size_t, [_i, _u32]),
    "pl_ntt_host": (_i, [_vp, _vp, _i, _u32, _i, _vp, _i, _vp, _vp]),
    "pl_ntt_host_groups": (_i, [_i]),
    "pl_ntt_host_groups_set": (_i, [_i, _i]),
    "pl_lde_domain_create": (_i, [_i, _i, _vp, ctypes.POINTER(_vp)]),
    "pl_lde_domain_destroy": (_i, [_vp]),
    "pl_lde_domain_bytes": (ctypes.c_size_t, [_vp]),
    "pl_lde_work_bytes": (ctypes.c_size_t, [_i, _i]),
    "pl_lde": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "pl_lde3_domain_create": (_i, [_i, _vp, ctypes.POINTER(_vp)]),
    "pl_lde_domain_work_bytes": (ctypes.c_size_t, [_vp]),
    "pl_lde_domain_blowup": (_i, [_vp]),
    "pl_coset_domain_create": (_i, [_i, _vp, ctypes.POINTER(_vp)]),
    "pl_coset_domain_destroy": (_i, [_vp]),
    "pl_coset_domain_bytes": (ctypes.c_size_t, [_vp]),
    "pl_coset_work_bytes": (ctypes.c_size_t, [_i]),
    "pl_coset_ntt": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "pl_coset_scale": (_i, [_vp, _u64, _u64, _i, _vp, _vp]),
    "pl_coset_ntt_step": (_i, [_vp, _vp, _i, _i, _vp, _vp]),
    "pl_scan_work_bytes": (ctypes.c_size_t, [_u64]),
    "pl_grand_product": (_i, [_vp, _u64, _vp, _vp, _vp]),
    "pl_grand_product_noalloc": (_i, [_vp, _u64, _vp, _vp]),
    "pl_eval": (_i, [_vp, _u64, _vp, _vp, _vp, _vp]),
    "pl_div_linear": (_i, [_vp, _u64, _vp, _vp, _vp, _vp]),
    "pl_vec_work_bytes": (ctypes.c_size_t, [_u64]),
    "pl_ntt_naive": (_i, [_vp, _vp, _i, _vp, _vp]),
    "pl_lde_naive": (_i, [_vp, _vp, _i, _i, _vp, _vp]),
    "pl_vec_op": (_i, [_i, _vp, _vp, _vp, _u64, _vp, _vp]),
    "pl_exp_mod": (_i, [_vp, _vp, _u64, _i, _vp]),
    "pl_gen_sequence": (_i, [_u64, _vp, _vp, _vp]),
    "pl_repeat_to_poly": (_i, [_vp, _vp, _u64, _vp]),
    "pl_perm_ratio": (_i, [ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "pl_perm_grand_product": (_i, [ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp, _vp, _vp, _i, _vp, _vp, _vp,
                                   _vp]),
    "pl_batch_create": (_i, [_vp, _i, _i, ctypes.POINTER(_vp)]),
    "pl_batch_run": (_i, [_vp, _vp]),
    "pl_batch_run_plain": (_i, [_vp, _vp]),
    "pl_batch_info": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_size_t)]),
    "pl_batch_destroy": (None, [_vp]),
    "pl_server_create": (_i, [_vp, ctypes.POINTER(_vp)]),
    "pl_server_submit": (_i, [_vp, _vp, ctypes.POINTER(_u64)]),
    "pl_server_wait": (_i, [_vp, _u64, _vp]),
    "pl_server_release": (_i, [_vp, _u64]),
    "pl_server_get_stats": (_i, [_vp, _vp]),
    "pl_server_now": (ctypes.c_double, [_vp]),
    "pl_job_bytes": (ctypes.c_size_t, [_i, _i]),
    "pl_server_destroy": (None, [_vp]),
    "pl_uniform_work_bytes": (ctypes.c_size_t, [_u64]),
    "pl_fill_uniform": (_i, [_vp, _u64, _u64, _vp, _vp]),
    "pl_digest": (_i, [_vp, _u64, _vp, _vp]),
    "pl_probe_mad_rate": (_i, [ctypes.POINTER(ctypes.c_double), _vp]),
    "pl_autotune": (_i, [_i, _i, ctypes.c_char_p, _i, _i]),
    "pl_plan_get": (_i, [_i, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "pl_plan_set": (_i, [_i, ctypes.POINTER(_i), _i]),
    "pl_plan_load": (_i, [ctypes.c_char_p]),
    "pl_plan_candidates": (_i, [_i, ctypes.POINTER(_i), _i]),
}

SERVE_BATCHED, SERVE_PLAIN = 0, 1


class ServerConfig(ctypes.Structure):
    """pl_server_config"""
    _fields_ = [("hbm_bytes", ctypes.c_size_t), ("slots", _i), ("mode", _i), ("max_group_jobs", _i),
                ("group_elems", _u64), ("large_work", _u64)]


class JobDesc(ctypes.Structure):
    """pl_job_desc"""
    _fields_ = [("lg", _i), ("lg_blowup", _i), ("d_evals", _vp * 8), ("d_beta", _vp), ("d_gamma", _vp),
                ("d_zeta", _vp)]


class JobResult(ctypes.Structure):
    """pl_job_result"""
    _fields_ = [("status", _i), ("arrival_s", ctypes.c_double), ("admit_s", ctypes.c_double),
                ("done_s", ctypes.c_double), ("d_coeffs", _vp), ("d_lde", _vp), ("d_z", _vp), ("d_y", _vp),
                ("d_q", _vp)]


class ServerStats(ctypes.Structure):
    """pl_server_stats"""
    _fields_ = [("submitted", _u64), ("completed", _u64), ("groups", _u64), ("queued", _u64),
                ("arena_bytes", ctypes.c_size_t), ("arena_peak", ctypes.c_size_t),
                ("arena_used", ctypes.c_size_t), ("domain_bytes", ctypes.c_size_t)]

BATCH_OPS = {"ntt": 0, "intt": 1, "lde": 2, "eval_div": 3}
PL_BATCH_GRAPH = 1


class BatchItem(ctypes.Structure):
    """pl_batch_item"""
    _fields_ = [("op", _i), ("lg", _i), ("lg_blowup", _i), ("d_in", _vp), ("d_out", _vp), ("d_z", _vp),
                ("d_y", _vp)]

VEC_OPS = {"add": 0, "sub": 1, "mul": 2, "div": 3, "add_scalar": 4, "sub_scalar": 5, "mul_scalar": 6,
           "div_scalar": 7, "neg": 8, "inv": 9, "to_mont": 10, "from_mont": 11}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libpl.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(fn, status):
    if status != PL_OK:
        raise PlError(fn, status)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class Domain:
    """Twiddle tables for NTT/iNTT of size 2^lg (the reference's Ntt/Intt Params)."""

    def __init__(self, lg):
        self.lg = lg
        h = ctypes.c_void_p()
        check("pl_domain_create", lib().pl_domain_create(lg, ctypes.byref(h)))
        self.h = h

    @property
    def nbytes(self):
        return lib().pl_domain_bytes(self.h)

    def close(self):
        if self.h:
            lib().pl_domain_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ntt(x, dom, inverse=False, out=None, work=None, stream=None):
    """NTT (or iNTT) of x, shape (2^lg, 4) int64 on the GPU.  out may be x."""
    import torch
    n = 1 << dom.lg
    assert x.shape == (n, 4) and x.dtype == torch.int64 and x.is_cuda
    if out is None:
        out = torch.empty_like(x)
    check("pl_ntt", lib().pl_ntt(_ptr(x), _ptr(out), dom.lg, 1 if inverse else 0, dom.h, _ptr(work),
                                 _stream(stream)))
    return out


def ntt_batch(x, dom, npoly, inverse=False, out=None, work=None, stream=None):
    import torch
    n = 1 << dom.lg
    assert x.shape == (npoly * n, 4) and x.dtype == torch.int64 and x.is_cuda
    if out is None:
        out = torch.empty_like(x)
    check("pl_ntt_batch", lib().pl_ntt_batch(_ptr(x), _ptr(out), dom.lg, npoly, 1 if inverse else 0, dom.h,
                                             _ptr(work), _stream(stream)))
    return out


class LdeDomain:
    """Coset tables for our_lde: 2^lg_n coefficients -> 2^(lg_n+lg_blowup) values.
    shift: None (the reference's 7^(2^lg_blowup)) or a numpy (1, 4) uint64 Montgomery element."""

    def __init__(self, lg_n, lg_blowup, shift=None, blowup3=False):
        """blowup3: the blowup-3 domain (cosets shift * w_3N^t; lg_blowup ignored)"""
        import numpy as np
        self.lg_n, self.lg_blowup = lg_n, lg_blowup
        h = ctypes.c_void_p()
        sp = None
        if shift is not None:
            self._shift = np.ascontiguousarray(shift, dtype=np.uint64).reshape(4)
            sp = ctypes.c_void_p(self._shift.ctypes.data)
        if blowup3:
            check("pl_lde3_domain_create", lib().pl_lde3_domain_create(lg_n, sp, ctypes.byref(h)))
        else:
            check("pl_lde_domain_create", lib().pl_lde_domain_create(lg_n, lg_blowup, sp, ctypes.byref(h)))
        self.h = h
        self.blowup = lib().pl_lde_domain_blowup(h)
        self.work_bytes = lib().pl_lde_domain_work_bytes(h)

    @property
    def nbytes(self):
        return lib().pl_lde_domain_bytes(self.h)

    def close(self):
        if self.h:
            lib().pl_lde_domain_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lde(c, dom, out=None, work=None, stream=None):
    import torch
    n = 1 << dom.lg_n
    assert c.shape == (n, 4) and c.dtype == torch.int64 and c.is_cuda
    if out is None:
        out = torch.empty((n * dom.blowup, 4), dtype=torch.int64, device=c.device)
    check("pl_lde", lib().pl_lde(_ptr(c), _ptr(out), dom.h, _ptr(work), _stream(stream)))
    return out


class CosetDomain:
    """Tables of the coset transform over g<w_N> (Ntt_coset / Intt_coset);
    shift: None (g = 7) or a numpy (1, 4) uint64 Montgomery element."""

    def __init__(self, lg, shift=None):
        import numpy as np
        self.lg = lg
        sp = None
        if shift is not None:
            self._shift = np.ascontiguousarray(shift, dtype=np.uint64).reshape(4)
            sp = ctypes.c_void_p(self._shift.ctypes.data)
        h = ctypes.c_void_p()
        check("pl_coset_domain_create", lib().pl_coset_domain_create(lg, sp, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().pl_coset_domain_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def coset_ntt(x, dom, inverse=False, out=None, work=None, stream=None):
    import torch
    if out is None:
        out = torch.empty_like(x)
    check("pl_coset_ntt", lib().pl_coset_ntt(_ptr(x), _ptr(out), 1 if inverse else 0, dom.h, _ptr(work),
                                             _stream(stream)))
    return out


def coset_scale(chunk, chunk_id, dom, inverse=False, stream=None):
    check("pl_coset_scale", lib().pl_coset_scale(_ptr(chunk), chunk.shape[0], chunk_id, 1 if inverse else 0, dom.h,
                                                 _stream(stream)))
    return chunk


def coset_ntt_step(x, out, step, dom, inverse=False, stream=None):
    check("pl_coset_ntt_step", lib().pl_coset_ntt_step(_ptr(x), _ptr(out), step, 1 if inverse else 0, dom.h,
                                                       _stream(stream)))
    return out


def grand_product(p, incl_out=None, work=None, stream=None):
    """in place: p[i] <- prod_{j<i} p[j]"""
    check("pl_grand_product", lib().pl_grand_product(_ptr(p), p.shape[0], _ptr(incl_out), _ptr(work),
                                                     _stream(stream)))
    return p


def grand_product_noalloc(p, incl_out=None, stream=None):
    """in place: p[i] <- prod_{j<i} p[j], library-held scratch (no allocation)"""
    check("pl_grand_product_noalloc", lib().pl_grand_product_noalloc(_ptr(p), p.shape[0], _ptr(incl_out),
                                                                     _stream(stream)))
    return p


def evaluate(poly, x, y=None, work=None, stream=None):
    import torch
    if y is None:
        y = torch.empty((1, 4), dtype=torch.int64, device=x.device)
    check("pl_eval", lib().pl_eval(_ptr(poly), poly.shape[0], _ptr(x), _ptr(y), _ptr(work), _stream(stream)))
    return y


def div_linear(a, c, rem=None, work=None, stream=None):
    """in place: a <- quotient of a(X) by (X + c)"""
    check("pl_div_linear", lib().pl_div_linear(_ptr(a), a.shape[0], _ptr(c), _ptr(rem), _ptr(work),
                                               _stream(stream)))
    return a


def vec_op(op, a, b=None, out=None, work=None, stream=None):
    import torch
    if out is None:
        out = torch.empty_like(a)
    check("pl_vec_op", lib().pl_vec_op(VEC_OPS[op], _ptr(a), _ptr(b), _ptr(out), a.shape[0], _ptr(work),
                                       _stream(stream)))
    return out


def exp_mod(a, e, out=None, stream=None):
    """out[i] = a[i]^e (e = 0: one, e = 1: copy)"""
    import torch
    if out is None:
        out = torch.empty_like(a)
    check("pl_exp_mod", lib().pl_exp_mod(_ptr(a), _ptr(out), a.shape[0], e, _stream(stream)))
    return out


def gen_sequence(n, x, out=None, stream=None):
    """out[i] = x^i, i < n"""
    import torch
    if out is None:
        out = torch.empty((n, 4), dtype=torch.int64, device=x.device)
    check("pl_gen_sequence", lib().pl_gen_sequence(n, _ptr(x), _ptr(out), _stream(stream)))
    return out


def repeat_to_poly(x, n, out=None, stream=None):
    """out[i] = x, i < n"""
    import torch
    if out is None:
        out = torch.empty((n, 4), dtype=torch.int64, device=x.device)
    check("pl_repeat_to_poly", lib().pl_repeat_to_poly(_ptr(x), _ptr(out), n, _stream(stream)))
    return out


def _ptr_array(ts):
    arr = (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])
    return ctypes.cast(arr, ctypes.POINTER(ctypes.c_void_p)), arr


def perm_ratio(w, s, beta, gamma, k, dom, out=None, grand_product=False, work=None, stream=None):
    """w, s: three device columns each; k: numpy (3, 4) uint64 Montgomery constants"""
    import numpy as np
    import torch
    if out is None:
        out = torch.empty_like(w[0])
    kk = np.ascontiguousarray(k, dtype=np.uint64).reshape(3, 4)
    wp, _w = _ptr_array(w)
    sp, _s = _ptr_array(s)
    fn = lib().pl_perm_grand_product if grand_product else lib().pl_perm_ratio
    check("pl_perm_ratio", fn(wp, sp, _ptr(beta), _ptr(gamma), ctypes.c_void_p(kk.ctypes.data), dom.lg, dom.h,
                              _ptr(out), _ptr(work), _stream(stream)))
    return out


class Batch:
    """A planned batch over caller-owned tensors.

    items: list of dicts {op: "ntt"|"intt"|"lde"|"eval_div", lg, lg_blowup,
    inp, out, z, y} (tensors on the current device).  The tensors are kept
    referenced for the batch's lifetime."""

    def __init__(self, items, graph=True):
        arr = (BatchItem * len(items))()
        self._keep = []
        for i, it in enumerate(items):
            arr[i].op = BATCH_OPS[it["op"]]
            arr[i].lg = it["lg"]
            arr[i].lg_blowup = it.get("lg_blowup", 0)
            arr[i].d_in = it["inp"].data_ptr()
            arr[i].d_out = it["out"].data_ptr()
            arr[i].d_z = it["z"].data_ptr() if it.get("z") is not None else None
            arr[i].d_y = it["y"].data_ptr() if it.get("y") is not None else None
            self._keep.append(it)
        self.h = ctypes.c_void_p()
        check("pl_batch_create", lib().pl_batch_create(ctypes.cast(arr, _vp), len(items),
                                                       PL_BATCH_GRAPH if graph else 0, ctypes.byref(self.h)))

    def run(self, stream=None):
        check("pl_batch_run", lib().pl_batch_run(self.h, _stream(stream)))

    def run_plain(self, stream=None):
        check("pl_batch_run_plain", lib().pl_batch_run_plain(self.h, _stream(stream)))

    def info(self):
        n = ctypes.c_int()
        b = ctypes.c_size_t()
        check("pl_batch_info", lib().pl_batch_info(self.h, ctypes.byref(n), ctypes.byref(b)))
        return n.value, b.value

    def close(self):
        if self.h:
            lib().pl_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop for proof jobs (pl_server_*).  Inputs are device tensors
    kept referenced until the job is waited for."""

    def __init__(self, hbm_bytes=0, slots=0, mode=SERVE_BATCHED, max_group_jobs=0, group_elems=0, large_work=0):
        cfg = ServerConfig(hbm_bytes, slots, mode, max_group_jobs, group_elems, large_work)
        self.h = ctypes.c_void_p()
        self._keep = {}
        check("pl_server_create", lib().pl_server_create(ctypes.byref(cfg), ctypes.byref(self.h)))

    def submit(self, lg, lg_blowup, evals, beta, gamma, zeta):
        d = JobDesc()
        d.lg, d.lg_blowup = lg, lg_blowup
        for c in range(8):
            d.d_evals[c] = evals[c].data_ptr()
        d.d_beta, d.d_gamma, d.d_zeta = beta.data_ptr(), gamma.data_ptr(), zeta.data_ptr()
        jid = ctypes.c_uint64()
        check("pl_server_submit", lib().pl_server_submit(self.h, ctypes.byref(d), ctypes.byref(jid)))
        self._keep[jid.value] = (evals, beta, gamma, zeta)
        return jid.value

    def wait(self, jid):
        r = JobResult()
        st = lib().pl_server_wait(self.h, jid, ctypes.byref(r))
        self._keep.pop(jid, None)
        if st != PL_OK:
            raise PlError("pl_server_wait", st)
        return r

    def release(self, jid):
        check("pl_server_release", lib().pl_server_release(self.h, jid))

    def now(self):
        return lib().pl_server_now(self.h)

    def stats(self):
        st = ServerStats()
        check("pl_server_get_stats", lib().pl_server_get_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            lib().pl_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def probe_mad_rate(stream=None):
    """Measured v_mad_u64_u32 peak of the device, 10^12 per second."""
    v = ctypes.c_double()
    check("pl_probe_mad_rate", lib().pl_probe_mad_rate(ctypes.byref(v), _stream(stream)))
    return v.value


def autotune(lg_lo, lg_hi, cache_path=None, reps=5, remeasure=False, host=False):
    """pass splits of each size (and, host=True, the host-to-host copy groups)"""
    check("pl_autotune", lib().pl_autotune(lg_lo, lg_hi, cache_path.encode() if cache_path else None, reps,
                                           (1 if remeasure else 0) | (2 if host else 0)))


def host_groups(lg):
    """lg_k that pl_ntt_host uses for lg_k < 0 (tuned, else the default)"""
    v = lib().pl_ntt_host_groups(lg)
    check("pl_ntt_host_groups", min(v, 0))
    return v


def plan_get(lg):
    m = (_i * 8)()
    P = _i()
    check("pl_plan_get", lib().pl_plan_get(lg, m, ctypes.byref(P)))
    return [m[i] for i in range(P.value)]


def plan_set(lg, m):
    arr = (_i * 8)(*m) if m else None
    check("pl_plan_set", lib().pl_plan_set(lg, arr, len(m) if m else 0))


def plan_candidates(lg):
    n = lib().pl_plan_candidates(lg, None, 0)
    buf = (_i * (8 * max(n, 1)))()
    lib().pl_plan_candidates(lg, buf, n)
    return [[buf[8 * i + k] for k in range(8) if buf[8 * i + k]] for i in range(n)]


def job_bytes(lg, lg_blowup):
    return lib().pl_job_bytes(lg, lg_blowup)


def fill_uniform(out, seed, stream=None):
    """out: device tensor (n, 4) int64; the seeded stream of synth.uniform."""
    check("pl_fill_uniform", lib().pl_fill_uniform(_ptr(out), out.shape[0], seed & 0xFFFFFFFFFFFFFFFF, None,
                                                   _stream(stream)))
    return out


def digest_ptr(ptr, n, stream=None):
    """64-bit digest of n elements at a device address (see synth.digest)."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    with torch.cuda.stream(s):
        o = torch.zeros(1, dtype=torch.int64, device="cuda")
        check("pl_digest", lib().pl_digest(ctypes.c_void_p(ptr), n, _ptr(o), _stream(s)))
        s.synchronize()
        return int(o.cpu().numpy().view("uint64")[0])


def digest(t, stream=None):
    return digest_ptr(t.data_ptr(), t.shape[0], stream)


def copy_from_ptr(ptr, n, stream=None):
    """Device copy of n elements at a raw device address into a new tensor
    (out = a + 0 through the element-wise operator)."""
    import torch
    out = torch.empty((n, 4), dtype=torch.int64, device="cuda")
    zero = torch.zeros((1, 4), dtype=torch.int64, device="cuda")
    check("pl_vec_op", lib().pl_vec_op(VEC_OPS["add_scalar"], ctypes.c_void_p(ptr), _ptr(zero), _ptr(out), n, None,
                                       _stream(stream)))
    return out


def ntt_pass(x, out, dom, p, npoly=1, inverse=False, stream=None):
    check("pl_ntt_pass", lib().pl_ntt_pass(_ptr(x), _ptr(out), dom.lg, npoly, p, 1 if inverse else 0, dom.h,
                                           _stream(stream)))
    return out


def params(stream=None):
    """The reference-layout Params buffer (4 x 16384 elements) on the device."""
    import torch
    out = torch.empty((4 * 16384, 4), dtype=torch.int64, device="cuda")
    check("pl_params_fill", lib().pl_params_fill(_ptr(out), _stream(stream)))
    return out


def stage_no_rotate(chunk, lg_domain, iterations, stage, chunk_id, params_buf, stream=None):
    """One Ntt::stage_no_rotate call on a device chunk (n x 4 int64 tensor), in place."""
    check("pl_ntt_stage_no_rotate",
          lib().pl_ntt_stage_no_rotate(_ptr(chunk), chunk.shape[0], lg_domain, iterations, stage, chunk_id,
                                       _ptr(params_buf), _stream(stream)))
    return chunk


def ct_stage_indices(tid, stage, iterations):
    """The chunked flow's element pair of thread tid (ntt.cu:168-178)."""
    a, b = _u64(), _u64()
    lib().pl_ct_stage_indices(tid, stage, iterations, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def rotate_stage_output_index(g, stage, iterations):
    return lib().pl_rotate_stage_output_index(g, stage, iterations)


def stage_gather(src, chunk, chunk_id, stage, iterations, stream=None):
    check("pl_stage_gather", lib().pl_stage_gather(_ptr(src), _ptr(chunk), chunk.shape[0], chunk_id, stage,
                                                   iterations, _stream(stream)))
    return chunk


def stage_scatter(chunk, dst, chunk_id, stage, iterations, chunk_local=False, stream=None):
    check("pl_stage_scatter", lib().pl_stage_scatter(_ptr(chunk), _ptr(dst), chunk.shape[0], chunk_id, stage,
                                                     iterations, 1 if chunk_local else 0, _stream(stream)))
    return dst


def stage_transition(src, dst, chunk_id, prev_stage, prev_it, next_stage, next_it, stream=None):
    check("pl_stage_transition", lib().pl_stage_transition(_ptr(src), _ptr(dst), src.shape[0], chunk_id, prev_stage,
                                                           prev_it, next_stage, next_it, _stream(stream)))
    return dst


def bit_reverse(x, out=None, stream=None):
    import torch
    if out is None:
        out = torch.empty_like(x)
    lg = x.shape[0].bit_length() - 1
    check("pl_bit_reverse", lib().pl_bit_reverse(_ptr(x), _ptr(out), lg, _stream(stream)))
    return out


def ntt_segmented(x, lg_i, lg_k, params_buf, stream=None):
    """The reference's chunked flow on the device (forward NTT, in place)."""
    lg = x.shape[0].bit_length() - 1
    check("pl_ntt_segmented", lib().pl_ntt_segmented(_ptr(x), lg, lg_i, lg_k, _ptr(params_buf), None,
                                                     _stream(stream)))
    return x


def ntt_host(h, dom, lg_k=-1, npoly=1, inverse=False, out=None, work=None, stream=None):
    """Host-to-host NTT of npoly polynomials held in a host tensor (pinned for
    overlap), shape (npoly * 2^lg, 4) int64; returns the output host tensor.
    lg_k: 2^lg_k copy groups, or -1 for the size's tuned (else default) count.
    work: None (the call allocates its device slots) or a device tensor of at
    least pl_ntt_host_work_bytes(lg, npoly) bytes."""
    import torch
    if out is None:
        out = torch.empty_like(h)
    check("pl_ntt_host", lib().pl_ntt_host(ctypes.c_void_p(h.data_ptr()), ctypes.c_void_p(out.data_ptr()), dom.lg,
                                           npoly, 1 if inverse else 0, dom.h, lg_k, _ptr(work), _stream(stream)))
    return out


def ntt_num_passes(lg):
    return lib().pl_ntt_num_passes(lg)


def domain_num_passes(dom):
    return lib().pl_domain_num_passes(dom.h)


def ntt_naive(x, dom, out=None, stream=None):
    """The workload baseline NTT (bit reversal + one radix-2 launch per stage)."""
    import torch
    if out is None:
        out = torch.empty_like(x)
    check("pl_ntt_naive", lib().pl_ntt_naive(_ptr(x), _ptr(out), dom.lg, dom.h, _stream(stream)))
    return out


def lde_naive(c, dom, lg_blowup, out=None, stream=None):
    """naive_lde: out[k] = NTT_N(c)[k mod N] for k < N << lg_blowup (dom of size N)."""
    import torch
    if out is None:
        out = torch.empty((c.shape[0] << lg_blowup, 4), dtype=torch.int64, device=c.device)
    check("pl_lde_naive", lib().pl_lde_naive(_ptr(c), _ptr(out), dom.lg, lg_blowup, dom.h, _stream(stream)))
    return out


def ntt_work_bytes(lg, npoly=1):
    return lib().pl_ntt_work_bytes(lg, npoly)


def to_dev(a):
    """numpy uint64 (n, 4) -> torch int64 (n, 4) on cuda:0"""
    import numpy as np
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).to("cuda")


def to_host(t):
    import numpy as np
    return t.cpu().numpy().view(np.uint64)
