"""Host-side mirror of the reference inner-loop interface, over the C ABI.

Each method names the Fortran subroutine it stands in for
(cfengno1/Hofstadter_Hubbard_TSC, paths relative to the reference tree).
Device vectors are torch complex128 CUDA tensors (HIP on ROCm); all compute
runs in libhhdmrg.so on the context's stream.  There is no CPU fallback: every
call raises when the library or the GPU is missing.
"""
import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib as L


class HHError(RuntimeError):
    pass


def _check(ctx_ptr, rc, what):
    if rc != 0:
        lib = L.load()
        msg = lib.hh_last_error(ctx_ptr)
        raise HHError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))


def _dptr(t, n=None):
    """Device address of a contiguous complex128 torch tensor of at least n
    elements (the C side reads and writes n complex128 values through it)."""
    import torch
    if not hasattr(t, "is_cuda") or not t.is_cuda or not t.is_contiguous():
        raise HHError("expected a contiguous device tensor")
    if t.dtype != torch.complex128:
        raise HHError("expected a complex128 device tensor, got %s" % t.dtype)
    if n is not None and t.numel() < n:
        raise HHError("device tensor holds %d complex128 elements, %d needed" % (t.numel(), n))
    return C.c_void_p(t.data_ptr())


def _sync():
    import torch
    torch.cuda.synchronize()


def structure_factor_cut(cor, Lx, Ly, nx_cut=0):
    """Get_structure_factor_cut (code/measure.f90:2159-2236): the structure
    factor of a correlation matrix over the sites of columns nx_cut + 1 ..
    Lx - nx_cut (N' sites, site label y + (x-1) Ly as in Get_Lattice,
    code/warm_up.f90:796-812),

        S(kx, ky) = (1/N') [ sum_x C(x,x) + sum_{x != y} e^{i k.(r_x - r_y)} C(x,y) ],

    on the grid kx = 2 pi i / (Lx - 2 nx_cut), i = 0 .. Lx - 2 nx_cut, and
    ky = 2 pi j / Ly, j = 0 .. Ly, with C the real part of `cor` (the
    reference passes the correlation to a real array).  Returns (kx, ky, S),
    S real with shape (Lx - 2 nx_cut + 1, Ly + 1)."""
    c = np.real(np.asarray(cor))
    n = Lx * Ly
    if c.shape != (n, n):
        raise HHError("correlation matrix must be (Lx*Ly, Lx*Ly)")
    if nx_cut < 0 or 2 * nx_cut >= Lx:
        raise HHError("nx_cut leaves no column")
    kxn = Lx - 2 * nx_cut
    kx = 2.0 * np.pi * np.arange(kxn + 1) / kxn
    ky = 2.0 * np.pi * np.arange(Ly + 1) / Ly
    sites = np.arange(nx_cut * Ly, n - nx_cut * Ly)
    x = sites // Ly + 1.0
    y = sites % Ly + 1.0
    cs = c[np.ix_(sites, sites)]
    off = cs - np.diag(np.diag(cs))
    S = np.empty((kxn + 1, Ly + 1))
    for i, qx in enumerate(kx):
        for j, qy in enumerate(ky):
            ph = np.exp(1j * (qx * x + qy * y))
            S[i, j] = np.real(np.trace(cs) + ph @ off @ np.conj(ph))
    return kx, ky, S / len(sites)


class Batch:
    """Independent superblocks whose H.psi runs as one grouped launch per stage."""

    def __init__(self, ctx, ptr, sbs):
        self.ctx = ctx
        self.ptr = ptr
        self._keep = list(sbs)

    def __del__(self):
        try:
            if self.ptr:
                L.load().hh_batch_destroy(self.ptr)
        except Exception:
            pass

    def info(self):
        d, f = L.i64(), L.f64()
        L.load().hh_batch_info(self.ptr, C.byref(d), C.byref(f))
        return {"dim": d.value, "flops_alg": f.value}

    def set_graph(self, on):
        """Replay the batched Lanczos's grouped H.psi from a HIP graph (True)
        or launch its kernels one by one (False)."""
        L.load().hh_batch_set_graph(self.ptr, 1 if on else 0)


class Block:
    """A device-resident block (Total_Model, code/pubdata.f90:167-181)."""

    def __init__(self, ctx, ptr):
        self.ctx = ctx
        self.ptr = C.c_void_p(ptr) if not isinstance(ptr, C.c_void_p) else ptr
        self.ptr_owned = True

    def __del__(self):
        try:
            if self.ptr and self.ptr_owned:
                L.load().hh_block_destroy(self.ptr)
        except Exception:
            pass

    def info(self):
        lib = L.load()
        a, ns, nst, ln = L.i64(), L.i32(), L.i32(), L.i32()
        lib.hh_block_info(self.ptr, C.byref(a), C.byref(ns), C.byref(nst), C.byref(ln))
        return a.value, ns.value, nst.value, ln.value

    def sites(self):
        lib = L.load()
        _, _, nst, _ = self.info()
        out = np.zeros(nst, np.int32)
        lib.hh_block_sites(self.ptr, L.ip(out))
        return out

    def sectors(self):
        lib = L.load()
        _, ns, _, _ = self.info()
        N, S2, d = (np.zeros(ns, np.int32) for _ in range(3))
        h = np.zeros(ns, np.int64)
        lib.hh_block_sectors(self.ptr, L.ip(N), L.ip(S2), L.ip(d), L.ip(h))
        return N, S2, d, h

    def download(self):
        lib = L.load()
        a, _, _, _ = self.info()
        out = np.zeros(max(a, 1), np.complex128)
        _check(self.ctx.ptr, lib.hh_block_download(self.ctx.ptr, self.ptr, out.ctypes.data_as(C.c_void_p)),
               "hh_block_download")
        return out[:a]

    def upload(self, arena):
        """Whole arena from a complex128 numpy array or contiguous device tensor."""
        lib = L.load()
        size = self.info()[0]
        if hasattr(arena, "is_cuda") and arena.is_cuda:
            _sync()
            src = _dptr(arena, size)
        else:
            arena = np.ascontiguousarray(arena, dtype=np.complex128)
            if arena.size < size:
                raise HHError("arena holds %d complex128 elements, the block needs %d" % (arena.size, size))
            src = arena.ctypes.data_as(C.c_void_p)
        _check(self.ctx.ptr, lib.hh_block_upload(self.ctx.ptr, self.ptr, src), "hh_block_upload")


class Superblock:
    """hmerge + compiled model_wave tasks (code/mat_vec.f90:3476-3852)."""

    def __init__(self, ctx, ptr, sys, env):
        self.ctx = ctx
        self.ptr = ptr
        self._keep = (sys, env)

    def __del__(self):
        try:
            if self.ptr:
                L.load().hh_superblock_destroy(self.ptr)
        except Exception:
            pass

    def info(self):
        lib = L.load()
        d, n, fa, fe = L.i64(), L.i32(), L.f64(), L.f64()
        lib.hh_superblock_info(self.ptr, C.byref(d), C.byref(n), C.byref(fa), C.byref(fe))
        return {"dim": d.value, "npairs": n.value, "flops_alg": fa.value, "flops_exec": fe.value}

    @property
    def dim(self):
        return self.info()["dim"]

    def pairs(self):
        lib = L.load()
        n = self.info()["npairs"]
        s, e, r, c = (np.zeros(n, np.int32) for _ in range(4))
        o = np.zeros(n, np.int64)
        lib.hh_superblock_pairs(self.ptr, L.ip(s), L.ip(e), L.ip(r), L.ip(c), L.ip(o))
        return s, e, r, c, o

    def enlarged(self, side):
        lib = L.load()
        n = L.i32()
        lib.hh_superblock_enlarged(self.ptr, side, C.byref(n), None, None, None)
        N, S2, d = (np.zeros(n.value, np.int32) for _ in range(3))
        lib.hh_superblock_enlarged(self.ptr, side, C.byref(n), L.ip(N), L.ip(S2), L.ip(d))
        return N, S2, d


class Trunc:
    """Truncation operator (Get_truncation_operator, code/get_den.f90:184-428)."""

    def __init__(self, ctx, ptr, num_state=0, err=0.0, entropy=0.0):
        self.ctx = ctx
        self.ptr = ptr
        self.num_state = num_state
        self.err = err
        self.entropy = entropy

    def __del__(self):
        try:
            if self.ptr and getattr(self, "ptr_owned", True):
                L.load().hh_trunc_destroy(self.ptr)
        except Exception:
            pass

    def kept(self):
        lib = L.load()
        n = L.i32()
        lib.hh_trunc_kept(self.ptr, C.byref(n), None, None)
        a, k = np.zeros(n.value, np.int32), np.zeros(n.value, np.int32)
        lib.hh_trunc_kept(self.ptr, C.byref(n), L.ip(a), L.ip(k))
        return a, k

    def matrices(self):
        """[(enlarged sector, U)] in the new block's sector order (host copies)."""
        lib = L.load()
        a, k = self.kept()
        size = L.i64()
        off = np.zeros(len(a), np.int64)
        ld = np.zeros(len(a), np.int32)
        lib.hh_trunc_layout(self.ptr, C.byref(size), L.ip(off), L.ip(ld))
        arena = np.zeros(size.value, np.complex128)
        _check(self.ctx.ptr, lib.hh_trunc_download(self.ctx.ptr, self.ptr, arena.ctypes.data_as(C.c_void_p)),
               "hh_trunc_download")
        out = []
        for ai, ki, o, d in zip(a, k, off, ld):
            out.append((int(ai), arena[o:o + d * ki].reshape((d, ki), order="F").copy()))
        return out


@dataclass
class StepRecord:
    phase: int
    sys_len: int
    env_len: int
    iters: int
    num_state: int
    dim: int
    energy: float
    trunc_err: float
    entropy: float
    seconds: float
    matvec_seconds: float
    matvec_flops: float
    bind_seconds: float = 0.0
    wave_seconds: float = 0.0
    lanczos_seconds: float = 0.0
    trunc_seconds: float = 0.0
    grow_seconds: float = 0.0
    converged: bool = True
    ritz_change: float = 0.0
    residual: float = 0.0
    hpsi_fp64: int = 0
    hpsi_fp32: int = 0
    solver: int = 0
    fallback: bool = False

    @classmethod
    def from_c(cls, r):
        return cls(r.phase, r.sys_len, r.env_len, r.iters, r.num_state, r.dim, r.energy, r.trunc_err, r.entropy,
                   r.seconds, r.matvec_seconds, r.matvec_flops, r.bind_seconds, r.wave_seconds, r.lanczos_seconds,
                   r.trunc_seconds, r.grow_seconds, bool(r.converged), r.ritz_change, r.residual, r.hpsi_fp64,
                   r.hpsi_fp32, r.solver, bool(r.fallback))


class Context:
    """One GPU, one stream (hh_ctx)."""

    def __init__(self, device=0):
        lib = L.load()
        p = C.c_void_p()
        rc = lib.hh_ctx_create(device, C.byref(p))
        if rc != 0:
            raise HHError("hh_ctx_create failed (%d): %s" % (rc, lib.hh_last_error(None).decode()))
        self.ptr = p

    def close(self):
        if self.ptr:
            L.load().hh_ctx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _rc(self, rc, what):
        _check(self.ptr, rc, what)

    # -- model (code/quantum.f90:93-160, code/warm_up.f90:347-666) ---------------
    def set_model(self, model):
        self.model = model
        mp = L.model_params(model)
        self._rc(L.load().hh_set_model(self.ptr, C.byref(mp)), "hh_set_model")

    # -- blocks --------------------------------------------------------------
    def site_block(self, side):
        """Get_hamiltonian_site (code/warm_up.f90:1208-1244)."""
        p = C.c_void_p()
        self._rc(L.load().hh_block_site(self.ptr, side, C.byref(p)), "hh_block_site")
        return Block(self, p)

    def boundary_sites(self, length):
        """Sites of a block of `length` sites that keep operators (dmrg_setup0)."""
        n = L.i32()
        self._rc(L.load().hh_boundary_sites(self.ptr, length, C.byref(n), None), "hh_boundary_sites")
        out = np.zeros(n.value, np.int32)
        self._rc(L.load().hh_boundary_sites(self.ptr, length, C.byref(n), L.ip(out)), "hh_boundary_sites")
        return out.tolist()

    def create_block(self, side, length, sectors, sites):
        N = np.array([q[0] for q in sectors], np.int32)
        S2 = np.array([q[1] for q in sectors], np.int32)
        d = np.array([q[2] for q in sectors], np.int32)
        st = np.array(sites, np.int32)
        p = C.c_void_p()
        self._rc(L.load().hh_block_create(self.ptr, side, length, len(N), L.ip(N), L.ip(S2), L.ip(d), len(st),
                                          L.ip(st), C.byref(p)), "hh_block_create")
        return Block(self, p)

    def layout(self, block, kind):
        lib = L.load()
        n = L.i32()
        lib.hh_block_layout(block.ptr, kind, C.byref(n), None, None, None)
        bra, ket = np.zeros(n.value, np.int32), np.zeros(n.value, np.int32)
        off = np.zeros(n.value, np.int64)
        lib.hh_block_layout(block.ptr, kind, C.byref(n), L.ip(bra), L.ip(ket), L.ip(off))
        return bra, ket, off

    def op_offset(self, block, site_index, kind):
        o = L.i64()
        rc = L.load().hh_block_op_offset(block.ptr, site_index, kind, C.byref(o))
        return o.value if rc == 0 else -1

    # -- superblock ----------------------------------------------------------
    def bind(self, sys, env, N):
        """hmerge + compile model_wave (code/mat_vec.f90:3623-3852)."""
        p = C.c_void_p()
        self._rc(L.load().hh_superblock_bind(self.ptr, sys.ptr, env.ptr, N, C.byref(p)), "hh_superblock_bind")
        return Superblock(self, p, sys, env)

    def model_wave(self, sb, psi_in, psi_out):
        """model_wave (code/mat_vec.f90:3476): psi_out = H psi_in."""
        _sync()
        self._rc(L.load().hh_model_wave(self.ptr, sb.ptr, _dptr(psi_in, sb.dim), _dptr(psi_out, sb.dim)), "hh_model_wave")

    def batch(self, sbs):
        """Group bound superblocks for one-launch-per-stage H.psi (at most 32)."""
        arr = (C.c_void_p * len(sbs))(*[sb.ptr for sb in sbs])
        p = C.c_void_p()
        self._rc(L.load().hh_batch_create(self.ptr, len(sbs), arr, C.byref(p)), "hh_batch_create")
        return Batch(self, p, sbs)

    def batch_model_wave(self, batch, psi_in, psi_out):
        """psi_out[i] = H_i psi_in[i] for every member of the batch."""
        _sync()
        n = len(batch._keep)
        if len(psi_in) != n or len(psi_out) != n:
            raise HHError("one input and one output vector per batch member")
        a = (C.c_void_p * n)(*[_dptr(t, m.dim).value for t, m in zip(psi_in, batch._keep)])
        b = (C.c_void_p * n)(*[_dptr(t, m.dim).value for t, m in zip(psi_out, batch._keep)])
        self._rc(L.load().hh_batch_model_wave(self.ptr, batch.ptr, a, b), "hh_batch_model_wave")

    def set_autotune(self, on):
        """Tune the H.psi launch cut of every superblock bound and batch
        created from now on (cached by task-list shape statistics)."""
        self._rc(L.load().hh_set_autotune(self.ptr, 1 if on else 0), "hh_set_autotune")

    def set_launch_config(self, target_items=0, max_groups=32, edge=-1, band=4, wide=0):
        """Force one H.psi launch cut on every superblock bound and batch
        created from now on; target_items=0 clears it."""
        self._rc(L.load().hh_set_launch_config(self.ptr, int(target_items), int(max_groups), int(edge), int(band),
                                               int(wide)), "hh_set_launch_config")

    def autotune_stats(self):
        k, t, h, sec = L.i32(), L.i32(), L.i32(), L.f64()
        self._rc(L.load().hh_autotune_stats(self.ptr, C.byref(k), C.byref(t), C.byref(h), C.byref(sec)),
                 "hh_autotune_stats")
        return {"keys": k.value, "tunes": t.value, "hits": h.value, "seconds": sec.value}

    def batch_lanczos(self, batch, psis, lanzero, max_iter=100):
        """lanczos_memo_new of every member in lock step (one grouped H.psi per
        iteration); psis[i] updated in place.  Returns (e0, iters, converged)
        as numpy arrays."""
        _sync()
        n = len(batch._keep)
        if len(psis) != n:
            raise HHError("one start vector per batch member")
        a = (C.c_void_p * n)(*[_dptr(t, m.dim).value for t, m in zip(psis, batch._keep)])
        e0 = np.zeros(n)
        it = np.zeros(n, np.int32)
        cv = np.zeros(n, np.int32)
        self._rc(L.load().hh_batch_lanczos(self.ptr, batch.ptr, a, float(lanzero), int(max_iter), L.ip(e0), L.ip(it),
                                           L.ip(cv)), "hh_batch_lanczos")
        return e0, it, cv.astype(bool)

    def lanczos(self, sb, psi, lanzero, max_iter=100):
        """lanczos_memo_new (code/lanczos.f90:4-176), levels=1; psi updated in place."""
        _sync()
        e, it = L.f64(), L.i32()
        self._rc(L.load().hh_lanczos(self.ptr, sb.ptr, _dptr(psi, sb.dim), lanzero, max_iter, C.byref(e), C.byref(it)),
                 "hh_lanczos")
        return e.value, it.value

    def lanczos_levels(self, sb, psis, lanzero, max_iter=100):
        """lanczos_memo_new with levels = len(psis) (code/lanczos.f90:29-44):
        each device vector holds its level's start vector and receives its
        unit eigenvector; returns (energies, Lanczos vector counts)."""
        _sync()
        k = len(psis)
        ptrs = (C.c_void_p * k)(*[_dptr(p, sb.dim) for p in psis])
        e = np.zeros(k, np.float64)
        it = np.zeros(k, np.int32)
        self._rc(L.load().hh_lanczos_levels(self.ptr, sb.ptr, ptrs, k, lanzero, max_iter,
                                            e.ctypes.data_as(C.POINTER(C.c_double)),
                                            it.ctypes.data_as(C.POINTER(C.c_int32))), "hh_lanczos_levels")
        return e, it

    def lanczos_status(self):
        """(converged, last Ritz-value change, |H psi - E psi|) of the last lanczos call."""
        c, d, r = L.i32(), L.f64(), L.f64()
        self._rc(L.load().hh_lanczos_status(self.ptr, C.byref(c), C.byref(d), C.byref(r)), "hh_lanczos_status")
        return bool(c.value), d.value, r.value

    def density_matrix(self, sb, side, psi, psi2=None):
        """Get_density_matrix_sys_new / _env_new (code/get_den.f90:106-177):
        [(enlarged sector, rho)] in super-basis order, rho a host ndarray; with
        psi2 (a second level) the mixed (0.1 rho(psi) + 1.9 rho(psi2)) / 2."""
        import torch
        lib = L.load()
        nb, size = L.i32(), C.c_int64()
        self._rc(lib.hh_density_matrix_layout(sb.ptr, side, C.byref(nb), C.byref(size), None, None, None),
                 "hh_density_matrix_layout")
        sec = np.zeros(max(nb.value, 1), np.int32)
        dim = np.zeros(max(nb.value, 1), np.int32)
        off = np.zeros(max(nb.value, 1), np.int64)
        self._rc(lib.hh_density_matrix_layout(sb.ptr, side, C.byref(nb), C.byref(size),
                                              sec.ctypes.data_as(C.POINTER(C.c_int32)),
                                              dim.ctypes.data_as(C.POINTER(C.c_int32)),
                                              off.ctypes.data_as(C.POINTER(C.c_int64))), "hh_density_matrix_layout")
        rho = torch.zeros(max(size.value, 1), dtype=torch.complex128, device="cuda")
        _sync()
        if psi2 is None:
            self._rc(lib.hh_density_matrix(self.ptr, sb.ptr, side, _dptr(psi, sb.dim), _dptr(rho, size.value)),
                     "hh_density_matrix")
        else:
            self._rc(lib.hh_density_matrix_levels(self.ptr, sb.ptr, side, _dptr(psi, sb.dim), _dptr(psi2, sb.dim),
                                                  _dptr(rho, size.value)), "hh_density_matrix_levels")
        h = rho.cpu().numpy()
        return [(int(sec[i]), h[off[i]:off[i] + dim[i] * dim[i]].reshape((dim[i], dim[i]), order="F"))
                for i in range(nb.value)]

    def trunc_to_disk(self, trunc, length, path):
        """truns_to_disk (code/space_info.f90:1621-1636): the truncation that
        produces the block of `length` sites, in the reference's file layout."""
        self._rc(L.load().hh_trunc_to_disk(self.ptr, trunc.ptr, length, os.fsencode(path)), "hh_trunc_to_disk")

    def trunc_from_disk(self, prev, side, path):
        """truns_from_disk (code/space_info.f90:1642-1661) for the block grown
        from `prev`: (Trunc, length)."""
        p, n = C.c_void_p(), L.i32()
        self._rc(L.load().hh_trunc_from_disk(self.ptr, prev.ptr, side, os.fsencode(path), C.byref(n), C.byref(p)),
                 "hh_trunc_from_disk")
        return Trunc(self, p), n.value

    def basis_to_disk(self, block, path):
        """basis_to_disk (code/space_info.f90:1665-1706): Get_basis of `block`
        and one site as the reference's formatted text file."""
        self._rc(L.load().hh_basis_to_disk(self.ptr, block.ptr, os.fsencode(path)), "hh_basis_to_disk")

    def basis_from_disk(self, block, path):
        """basis_from_disk (code/space_info.f90:1710-1762): read the file and
        check it against Get_basis of `block`; returns the file's len."""
        n = L.i32()
        self._rc(L.load().hh_basis_from_disk(self.ptr, block.ptr, os.fsencode(path), C.byref(n)),
                 "hh_basis_from_disk")
        return n.value

    def wave_to_disk(self, sb, psi, path):
        """wave_to_disk (code/space_info.f90:1406-1436): psi in the reference's
        Fortran unformatted wavefunction file."""
        _sync()
        self._rc(L.load().hh_wave_to_disk(self.ptr, sb.ptr, _dptr(psi, sb.dim), os.fsencode(path)), "hh_wave_to_disk")

    def wave_from_disk(self, sb, path, psi):
        """wave_from_disk (code/space_info.f90:1440-1468) into the device vector psi."""
        _sync()
        self._rc(L.load().hh_wave_from_disk(self.ptr, sb.ptr, os.fsencode(path), _dptr(psi, sb.dim)), "hh_wave_from_disk")

    def truncate(self, sb, side, psi, kept_min, kept_max, trun_err, psi2=None):
        """Get_density_matrix_* + Get_truncation_operator (code/get_den.f90);
        with psi2 (a second level) on the mixed density matrix of two levels."""
        _sync()
        p = C.c_void_p()
        ns, er, en = L.i32(), L.f64(), L.f64()
        if psi2 is None:
            self._rc(L.load().hh_truncate(self.ptr, sb.ptr, side, _dptr(psi, sb.dim), kept_min, kept_max, trun_err,
                                          C.byref(p), C.byref(ns), C.byref(er), C.byref(en)), "hh_truncate")
        else:
            self._rc(L.load().hh_truncate_levels(self.ptr, sb.ptr, side, _dptr(psi, sb.dim), _dptr(psi2, sb.dim),
                                                 kept_min, kept_max, trun_err, C.byref(p), C.byref(ns), C.byref(er),
                                                 C.byref(en)), "hh_truncate_levels")
        return Trunc(self, p, ns.value, er.value, en.value)

    def trunc_identity(self, sb, side):
        p = C.c_void_p()
        self._rc(L.load().hh_trunc_identity(self.ptr, sb.ptr, side, C.byref(p)), "hh_trunc_identity")
        return Trunc(self, p)

    def grow(self, sb, side, trunc):
        """Get_Hamiltonian_Sys/Env + update_trun_* (code/get_ham.f90:4-659)."""
        p = C.c_void_p()
        self._rc(L.load().hh_grow_block(self.ptr, sb.ptr, side, trunc.ptr, C.byref(p)), "hh_grow_block")
        return Block(self, p)

    def renormalize(self, trunc, subs):
        """update_trun_dia / update_trun_ndia (code/update.f90:1387-1605) for one
        operator: subs = [(bra, ket, matrix)] in the truncation's enlarged
        basis (enlarged-sector indices, host complex matrices); returns
        [(kept bra, kept ket, U_bra^dag M U_ket)] for the pairs of kept sectors."""
        import torch
        lib = L.load()
        bra = np.array([b for b, _, _ in subs], np.int32)
        ket = np.array([k for _, k, _ in subs], np.int32)
        mats = [np.asarray(m, np.complex128) for _, _, m in subs]
        off = np.zeros(len(subs), np.int64)
        o = 0
        for i, m in enumerate(mats):
            off[i] = o
            o += m.size
        arena = np.concatenate([m.reshape(-1, order="F") for m in mats]) if mats else np.zeros(1, np.complex128)
        n, size = L.i32(), L.i64()
        self._rc(lib.hh_renormalize(self.ptr, trunc.ptr, len(subs), L.ip(bra), L.ip(ket), L.ip(off), None,
                                    C.byref(n), C.byref(size), None, None, None, None), "hh_renormalize")
        ob, ok = np.zeros(n.value, np.int32), np.zeros(n.value, np.int32)
        oo = np.zeros(n.value, np.int64)
        src = torch.from_numpy(arena).cuda()
        dst = torch.zeros(max(size.value, 1), dtype=torch.complex128, device="cuda")
        _sync()
        self._rc(lib.hh_renormalize(self.ptr, trunc.ptr, len(subs), L.ip(bra), L.ip(ket), L.ip(off),
                                    _dptr(src, arena.size), C.byref(n), C.byref(size), L.ip(ob), L.ip(ok), L.ip(oo),
                                    _dptr(dst, size.value)), "hh_renormalize")
        h = dst.cpu().numpy()
        keep = dict(zip(*trunc.kept()))
        sec = trunc.kept()[0]
        out = []
        for b_, k_, o_ in zip(ob, ok, oo):
            r, c = keep[sec[b_]], keep[sec[k_]]
            out.append((int(b_), int(k_), h[o_:o_ + r * c].reshape((r, c), order="F")))
        return out

    def conjugate_block(self, block):
        """Conjugate a block's arena in place (environment blocks taken from
        the reference's arrays, which hold the conjugated operators)."""
        self._rc(L.load().hh_block_conjugate(self.ptr, block.ptr), "hh_block_conjugate")

    def wave_trans(self, direction, old_sb, psi_old, sys_t, env_t, new_sb, psi_new):
        """wave_trans_LR (direction 0) / wave_trans_RL (1), code/warm_up.f90:1855-2170."""
        _sync()
        self._rc(L.load().hh_wave_trans(self.ptr, direction, old_sb.ptr, _dptr(psi_old, old_sb.dim), sys_t.ptr, env_t.ptr,
                                        new_sb.ptr, _dptr(psi_new, new_sb.dim)), "hh_wave_trans")

    def start_vector(self, psi, seed):
        _sync()
        self._rc(L.load().hh_start_vector(self.ptr, _dptr(psi), psi.numel(), seed), "hh_start_vector")

    # -- whole run (Get_Ground_State, code/quantum.f90:179-288) ---------------
    def dmrg_run(self, kept_max, n_sweeps, tol, kept_max_warm=0, max_iter=100, seed=1234,
                 warm_tol_factor=1000.0, first_tol_factor=15.0, precision=0, solver=0):
        """solver 0: plain FP64 Lanczos; 1: mixed-precision Lanczos rounds;
        2: the chosen Davidson form (hh_schedule.solver)."""
        s = L.Schedule(kept_max, kept_max_warm, n_sweeps, max_iter, tol, warm_tol_factor, first_tol_factor, seed,
                       precision, solver)
        cap = 4096
        recs = (L.StepRecord * cap)()
        n = L.i32()
        self._rc(L.load().hh_dmrg_run(self.ptr, C.byref(s), recs, cap, C.byref(n)), "hh_dmrg_run")
        return [StepRecord.from_c(r) for r in recs[:min(n.value, cap)]]

    # -- profiling -----------------------------------------------------------
    def set_profiling(self, on):
        L.load().hh_set_profiling(self.ptr, 1 if on else 0)

    def eigh_sectors(self, blocks, dims, keep, solver=0):
        """The truncation's eigensolver on given Hermitian blocks
        (hh_eigh_sectors): blocks is a complex128 device tensor holding the
        dims[i] x dims[i] column-major blocks one after another (overwritten:
        block i ends with the eigenvectors of its top keep[i] eigenvalues as
        its last keep[i] columns); returns the eigenvalues (device tensor,
        ascending per block)."""
        import torch
        dims = np.ascontiguousarray(dims, dtype=np.int32)
        keep = np.ascontiguousarray(keep, dtype=np.int32)
        total = int(np.sum(dims.astype(np.int64) ** 2))
        ptr = _dptr(blocks, total)
        w = torch.empty(int(dims.sum()), dtype=torch.float64, device=blocks.device)
        _sync()
        self._rc(L.load().hh_eigh_sectors(self.ptr, len(dims), L.ip(dims), L.ip(keep), ptr, C.c_void_p(w.data_ptr()),
                                          int(solver)), "hh_eigh_sectors")
        return w

    def set_eigensolver(self, solver):
        """0: hand-written density-matrix eigensolver (default); 1: rocSOLVER zheevd."""
        self._rc(L.load().hh_set_eigensolver(self.ptr, int(solver)), "hh_set_eigensolver")

    def eigensolver_status(self):
        """(solver, truncations handed back to zheevd, orthonormalisation
        steps and max |U^dag U - I| of the last hand-written call)."""
        a, b, c, d = L.i32(), L.i64(), L.i32(), L.f64()
        L.load().hh_eigensolver_status(self.ptr, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return a.value, b.value, c.value, d.value

    def set_precision(self, precision):
        """0: FP64 MFMA H.psi (plain form); 1: FP32 MFMA H.psi, FP64 vectors."""
        self._rc(L.load().hh_set_precision(self.ptr, int(precision)), "hh_set_precision")

    def stats(self):
        s = L.Stats()
        L.load().hh_get_stats(self.ptr, C.byref(s))
        return {f: getattr(s, f) for f, _ in L.Stats._fields_}

    def reset_stats(self):
        L.load().hh_reset_stats(self.ptr)


class Driver:
    """Step-wise sweep driver (hh_driver): warm_up_point, left_to_right and
    right_to_left of code/warm_up.f90:1284-1849, one call at a time, with all
    blocks and truncations resident on the device."""

    def __init__(self, ctx, kept_max, tol, kept_max_warm=0, max_iter=100, seed=1234, warm_tol_factor=1000.0,
                 first_tol_factor=15.0, precision=0, solver=0):
        self.ctx = ctx
        s = L.Schedule(kept_max, kept_max_warm, 0, max_iter, tol, warm_tol_factor, first_tol_factor, seed,
                       precision, solver)
        p = C.c_void_p()
        ctx._rc(L.load().hh_driver_create(ctx.ptr, C.byref(s), C.byref(p)), "hh_driver_create")
        self.ptr = p

    def __del__(self):
        try:
            if self.ptr:
                L.load().hh_driver_destroy(self.ptr)
        except Exception:
            pass

    @staticmethod
    def _recs(buf, n):
        return [StepRecord.from_c(r) for r in buf[:n]]

    def _psi(self, sys_len, psi):
        """Device address of psi after checking it against the dimension of
        the superblock of the stored sys block of sys_len sites and the env
        block of Nsite - sys_len - 2 sites (the C side reads that many
        complex128 values)."""
        n = self.ctx.model.Lx * self.ctx.model.Ly
        sysb, envb = self.block(0, sys_len), self.block(1, n - sys_len - 2)
        key = (sys_len, sysb.ptr.value, envb.ptr.value)
        if getattr(self, "_dims", None) is None:
            self._dims = {}
        if key not in self._dims:
            self._dims[key] = self.ctx.bind(sysb, envb, self.ctx.model.n_elec).dim
        return _dptr(psi, self._dims[key])

    def fork(self, window=None):
        """An independent copy of this driver's state (made on the device).
        window = ((sys_lo, sys_hi), (env_lo, env_hi)) copies only the blocks
        and truncations of those lengths (hh_driver_fork_window)."""
        p = C.c_void_p()
        if window is None:
            self.ctx._rc(L.load().hh_driver_fork(self.ptr, C.byref(p)), "hh_driver_fork")
        else:
            (slo, shi), (elo, ehi) = window
            self.ctx._rc(L.load().hh_driver_fork_window(self.ptr, int(slo), int(shi), int(elo), int(ehi), C.byref(p)),
                         "hh_driver_fork_window")
        d = Driver.__new__(Driver)
        d.ctx = self.ctx
        d.ptr = p
        return d

    def capture(self, max_samples):
        """The next max_samples plain (solver 0) steps record a training triple
        of the learned preconditioner each."""
        self.ctx._rc(L.load().hh_driver_capture(self.ptr, int(max_samples)), "hh_driver_capture")

    def train_precond(self, epochs=60, lr=1e-3):
        """Train the preconditioner network on the captured triples (Adam, on
        the GPU); returns (mean loss per epoch, number of triples)."""
        losses = np.zeros(max(int(epochs), 1))
        ns = L.i32()
        self.ctx._rc(L.load().hh_driver_train_precond(self.ptr, int(epochs), float(lr),
                                                      losses.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ns)),
                     "hh_driver_train_precond")
        return losses, ns.value

    def set_precond(self, mode):
        """0 the trained network, 1 the fixed diagonal form, 2 none (t = r)."""
        self.ctx._rc(L.load().hh_driver_set_precond(self.ptr, int(mode)), "hh_driver_set_precond")

    def set_min_dim(self, min_dim):
        """Solvers 1 and 2 hand superblocks below min_dim to the plain Lanczos
        (default 262144; 0 = never)."""
        self.ctx._rc(L.load().hh_driver_set_min_dim(self.ptr, int(min_dim)), "hh_driver_set_min_dim")

    def set_subspace(self, kmax):
        """Largest Davidson subspace of solver 2 before a restart (default 32,
        at most 128)."""
        self.ctx._rc(L.load().hh_driver_set_subspace(self.ptr, int(kmax)), "hh_driver_set_subspace")

    @property
    def levels(self):
        """Levels every step targets (1 unless set or restored by load)."""
        v = L.i32()
        self.ctx._rc(L.load().hh_driver_levels(self.ptr, C.byref(v)), "hh_driver_levels")
        return v.value

    def set_levels(self, levels):
        """Levels every later step targets (default 1, at most 8; plain
        Lanczos only): each step solves them one by one, projecting each off
        the ones below, and truncates on the mixed density matrix of levels 1
        and 2.  Records carry level 1's energy; level_energies() gives all."""
        self.ctx._rc(L.load().hh_driver_set_levels(self.ptr, int(levels)), "hh_driver_set_levels")

    def set_level_start(self, carry):
        """Start vectors of levels 2.. after a wave transformation: False
        seeded (the reference's), True the transformed previous levels."""
        self.ctx._rc(L.load().hh_driver_set_level_start(self.ptr, 1 if carry else 0), "hh_driver_set_level_start")

    def level_energies(self, max_rec=100000):
        """Level energies of the records the last warmup / sweep call
        returned, shape (records, levels)."""
        n = L.i32()
        self.ctx._rc(L.load().hh_driver_level_energies(self.ptr, 0, None, C.byref(n)), "hh_driver_level_energies")
        e = np.zeros((max(min(n.value, max_rec), 0), self.levels))
        self.ctx._rc(L.load().hh_driver_level_energies(self.ptr, e.shape[0], e.ctypes.data_as(C.POINTER(C.c_double)),
                                                       C.byref(n)), "hh_driver_level_energies")
        return e

    def copy_precond(self, src):
        """Take a device copy of src's trained network (same context)."""
        self.ctx._rc(L.load().hh_driver_copy_precond(self.ptr, src.ptr), "hh_driver_copy_precond")

    def set_fallback(self, factor):
        """Solver 2 hands the solve to the mixed Lanczos unless the FP64
        residual falls below factor x its value at the previous restart."""
        self.ctx._rc(L.load().hh_driver_set_fallback(self.ptr, float(factor)), "hh_driver_set_fallback")

    def set_solver(self, solver):
        """Ground-state solver of the following steps: 0 plain FP64 Lanczos, 1
        mixed-precision Lanczos rounds, 2 the chosen Davidson form."""
        self.ctx._rc(L.load().hh_driver_set_solver(self.ptr, int(solver)), "hh_driver_set_solver")

    def wave(self):
        """(sys_len, env_len, psi): the ground state of the last sweep step as a
        new device tensor, with the lengths of the blocks of its superblock."""
        import torch
        ls, le, d = L.i32(), L.i32(), L.i64()
        self.ctx._rc(L.load().hh_driver_wave(self.ptr, C.byref(ls), C.byref(le), C.byref(d), None), "hh_driver_wave")
        psi = torch.empty(d.value, dtype=torch.complex128, device="cuda")
        _sync()
        self.ctx._rc(L.load().hh_driver_wave(self.ptr, C.byref(ls), C.byref(le), C.byref(d), _dptr(psi, d.value)),
                     "hh_driver_wave")
        return ls.value, le.value, psi

    def warmup(self):
        cap = 4096
        buf = (L.StepRecord * cap)()
        n = L.i32()
        self.ctx._rc(L.load().hh_driver_warmup(self.ptr, buf, cap, C.byref(n)), "hh_driver_warmup")
        return self._recs(buf, min(n.value, cap))

    def sweep(self, direction, bg, end, lanzero, trun_err, kept_max, fresh=False, phase=1):
        cap = 4096
        buf = (L.StepRecord * cap)()
        n = L.i32()
        self.ctx._rc(L.load().hh_driver_sweep(self.ptr, direction, bg, end, lanzero, trun_err, kept_max,
                                              1 if fresh else 0, phase, buf, cap, C.byref(n)), "hh_driver_sweep")
        return self._recs(buf, min(n.value, cap))

    def reset_transform(self):
        """Forget the previous step, so the next sweep starts from its stored
        psi or a fresh vector instead of a wave transformation."""
        self.ctx._rc(L.load().hh_driver_reset_transform(self.ptr), "hh_driver_reset_transform")

    def trunc(self, side, length):
        """Borrowed handle of the stored truncation that made the block of `length`."""
        p = C.c_void_p()
        if L.load().hh_driver_trunc(self.ptr, side, length, C.byref(p)) != 0:
            raise HHError("no stored truncation of length %d" % length)
        t = Trunc(self.ctx, p)
        t.ptr_owned = False
        return t

    def density(self, sys_len, psi):
        """<n_i> per lattice position (index i-1) in the ground state psi of the
        superblock of the stored sys block of sys_len sites (measure_operator_dia)."""
        _sync()
        out = np.zeros(self.ctx.model.Lx * self.ctx.model.Ly)
        self.ctx._rc(L.load().hh_driver_density(self.ptr, sys_len, self._psi(sys_len, psi),
                                                out.ctypes.data_as(C.POINTER(C.c_double))), "hh_driver_density")
        return out

    def density_cross(self, sys_len, psi):
        """<n_i n_j> (nsite x nsite, NaN for pairs on one side) across the centre."""
        _sync()
        n = self.ctx.model.Lx * self.ctx.model.Ly
        out = np.zeros((n, n))
        self.ctx._rc(L.load().hh_driver_density_cross(self.ptr, sys_len, self._psi(sys_len, psi),
                                                      out.ctypes.data_as(C.POINTER(C.c_double))),
                     "hh_driver_density_cross")
        return out

    def density_row(self, sys_len, site, psi):
        """<n_site n_j> for every lattice position j (index j-1; NaN at j = site),
        site on the sys side of the cut (1 .. sys_len + 1)."""
        _sync()
        out = np.zeros(self.ctx.model.Lx * self.ctx.model.Ly)
        self.ctx._rc(L.load().hh_driver_density_row(self.ptr, sys_len, site, self._psi(sys_len, psi),
                                                    out.ctypes.data_as(C.POINTER(C.c_double))),
                     "hh_driver_density_row")
        return out

    def spin_cross(self, sys_len, psi):
        """<S_i . S_j> (nsite x nsite, NaN for pairs on one side) across the centre."""
        _sync()
        n = self.ctx.model.Lx * self.ctx.model.Ly
        out = np.zeros((n, n))
        self.ctx._rc(L.load().hh_driver_spin_cross(self.ptr, sys_len, self._psi(sys_len, psi),
                                                   out.ctypes.data_as(C.POINTER(C.c_double))),
                     "hh_driver_spin_cross")
        return out

    def hopping_cross(self, sys_len, psi):
        """sum_sigma <c^dag_i c_j> (nsite x nsite complex, NaN for pairs on one side)
        across the centre."""
        _sync()
        n = self.ctx.model.Lx * self.ctx.model.Ly
        out = np.zeros((n, n, 2))
        self.ctx._rc(L.load().hh_driver_hopping_cross(self.ptr, sys_len, self._psi(sys_len, psi),
                                                      out.ctypes.data_as(C.POINTER(C.c_double))),
                     "hh_driver_hopping_cross")
        return out[..., 0] + 1j * out[..., 1]

    def pair_cross(self, sys_len, psi, bonds):
        """<Delta^dag_a Delta_b> (len(bonds) x len(bonds) complex) for bonds a in the
        sys block and b in the env block (lattice positions); NaN elsewhere."""
        _sync()
        bonds = np.ascontiguousarray(np.asarray(bonds, dtype=np.int32).reshape(-1, 2))
        nb = len(bonds)
        out = np.zeros((nb, nb, 2))
        self.ctx._rc(L.load().hh_driver_pair_cross(self.ptr, sys_len, self._psi(sys_len, psi), nb,
                                                   bonds.ctypes.data_as(C.POINTER(C.c_int32)),
                                                   out.ctypes.data_as(C.POINTER(C.c_double))),
                     "hh_driver_pair_cross")
        return out[..., 0] + 1j * out[..., 1]

    def spin_row(self, sys_len, site, psi):
        """<S_site . S_j> for every lattice position j (index j-1; NaN at j = site),
        site on the sys side of the cut (1 .. sys_len + 1)."""
        _sync()
        out = np.zeros(self.ctx.model.Lx * self.ctx.model.Ly)
        self.ctx._rc(L.load().hh_driver_spin_row(self.ptr, sys_len, site, self._psi(sys_len, psi),
                                                 out.ctypes.data_as(C.POINTER(C.c_double))),
                     "hh_driver_spin_row")
        return out

    def hopping_row(self, sys_len, site, psi):
        """sum_sigma <c^dag_site c_j> for every lattice position j (complex, index
        j-1; NaN at j = site), site on the sys side (1 .. sys_len + 1)."""
        _sync()
        out = np.zeros((self.ctx.model.Lx * self.ctx.model.Ly, 2))
        self.ctx._rc(L.load().hh_driver_hopping_row(self.ptr, sys_len, site, self._psi(sys_len, psi),
                                                    out.ctypes.data_as(C.POINTER(C.c_double))),
                     "hh_driver_hopping_row")
        return out[:, 0] + 1j * out[:, 1]

    def _all(self, fn, sys_len, psi, cplx):
        _sync()
        n = self.ctx.model.Lx * self.ctx.model.Ly
        out = np.zeros((n, n, 2) if cplx else (n, n))
        self.ctx._rc(getattr(L.load(), fn)(self.ptr, sys_len, self._psi(sys_len, psi), out.ctypes.data_as(C.POINTER(C.c_double))),
                     fn)
        return out[:, :, 0] + 1j * out[:, :, 1] if cplx else out

    def density_all(self, sys_len, psi):
        """<n_i n_j> for every pair of lattice positions and <n_i^2> on the
        diagonal (Get_oper_cor_ndia1 with n): (Nsite, Nsite), symmetric."""
        return self._all("hh_driver_density_all", sys_len, psi, False)

    def spin_all(self, sys_len, psi):
        """<S_i . S_j> for every pair and <S_i . S_i> on the diagonal."""
        return self._all("hh_driver_spin_all", sys_len, psi, False)

    def hopping_all(self, sys_len, psi):
        """sum_sigma <c^dag_i c_j> for every pair and <n_i> on the diagonal
        (complex, Hermitian)."""
        return self._all("hh_driver_hopping_all", sys_len, psi, True)

    def save(self, path):
        """Write the driver's blocks, truncations, carried wave function and step
        counter to `path` (checkpoint)."""
        _sync()
        self.ctx._rc(L.load().hh_driver_save(self.ptr, os.fsencode(path)), "hh_driver_save")

    def load(self, path):
        """Replace the driver's state by the checkpoint at `path`."""
        _sync()
        self.ctx._rc(L.load().hh_driver_load(self.ptr, os.fsencode(path)), "hh_driver_load")

    def pair_block(self, sys_len, psi, quads):
        """<Delta^dag_a Delta_b> (complex) for each ((s1, s2), (s3, s4)) in quads:
        two bonds inside the sys block, every site of a before every site of b."""
        _sync()
        q = np.ascontiguousarray(np.asarray(quads, dtype=np.int32).reshape(-1, 4))
        out = np.zeros((len(q), 2))
        self.ctx._rc(L.load().hh_driver_pair_block(self.ptr, sys_len, self._psi(sys_len, psi), len(q),
                                                   q.ctypes.data_as(C.POINTER(C.c_int32)),
                                                   out.ctypes.data_as(C.POINTER(C.c_double))),
                     "hh_driver_pair_block")
        return out[:, 0] + 1j * out[:, 1]

    def block(self, side, length):
        """Borrowed handle of a stored block (do not destroy)."""
        p = C.c_void_p()
        rc = L.load().hh_driver_block(self.ptr, side, length, C.byref(p))
        if rc != 0:
            raise HHError("no stored block of length %d" % length)
        b = Block(self.ctx, p)
        b.ptr_owned = False
        return b
