"""Host-side mirror of the reference multigrid call surface over the C ABI.

``FieldSolver`` keeps the names and argument meaning of the reference
routines it stands in for:

=====================  ====================================================
``perform_vcycle``      mgridlib/mgridsteps.f90:384 (perform_vcycle)
``solve``               the equation modules' V-cycle loop
                        (plasma_pde/poissoneq.f90:427-450,
                        generic_pde/pdedefs.f90:378-402)
``find_AX``             mgridlib/discretepde.f90:672 (find_AX)
``find_bvec``           mgridlib/discretepde.f90:382 (find_bvec)
``smooth``              mgridlib/discretepde.f90:132 (perform_gseidel_smoothing)
``fill_ghosts``         mgridlib/discretepde.f90:1121 (update_boundary_ghostvalues)
``gmres``               mgridlib/gmres.f90:197 (performgmres)
=====================  ====================================================

Arrays are the reference's: ghosted Fortran-order float64 arrays of shape
(nx+2, ny+2, nz+2) for solutions and coefficients, interior (nx, ny, nz)
arrays for mgridsterm and operator outputs. A numpy array is host memory;
a torch tensor on a HIP device is used in place (its memory must be laid
out as the Fortran array, i.e. a C-contiguous tensor of shape
(nz+2, ny+2, nx+2) -- see :func:`to_device` / :func:`from_device`).

Every compute call goes to the HIP library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as M

_NAMES = {"ref": M.HIER_REF, "aligned": M.HIER_ALIGNED}
_PREC = {"fp64": M.PREC_FP64, "fp32dc": M.PREC_FP32_DC, "fp32_dc": M.PREC_FP32_DC}
_FLUX = {"upwind": M.FLUX_UPWIND, "sg": M.FLUX_SG}
_TOL = {"rms": M.TOL_RMS, "rel_l2": M.TOL_REL_L2, "rel": M.TOL_REL_L2}


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def ghosted(interior):
    """Zero-ghost Fortran array of shape (nx+2, ny+2, nz+2) around an interior array."""
    a = np.asarray(interior, dtype=np.float64)
    g = np.zeros((a.shape[0] + 2, a.shape[1] + 2, a.shape[2] + 2), order="F")
    g[1:-1, 1:-1, 1:-1] = a
    return g


def to_device(a, device="cuda"):
    """Fortran-order numpy array -> C-contiguous torch tensor with the same memory layout."""
    import torch
    a = np.asfortranarray(np.asarray(a, dtype=np.float64))
    return torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0))).to(device)


def from_device(t):
    """Inverse of :func:`to_device`."""
    return np.asfortranarray(t.detach().cpu().numpy().transpose(2, 1, 0))


STREAM_KINDS = {"read": 0, "write": 1, "copy": 2, "triad": 3, "sweep": 4}


def time_stream(kind, bytes_per_array=1 << 30, reps=10):
    """HBM rate (GB/s) of a hand-written streaming kernel with 16-byte accesses
    on the current device: kind "read", "write", "copy", "triad" (o = a + s b)
    or "sweep" (o = a + s (b + c): three streams read and one written, the
    red-black sweep's pattern). Returns (GB/s, mean ms per launch, bytes per
    launch)."""
    L = M.load()
    ms, by = C.c_double(), C.c_double()
    M.check(L.mps_time_stream(STREAM_KINDS[kind], int(bytes_per_array), int(reps), C.byref(ms), C.byref(by)))
    return by.value / (ms.value * 1e-3) / 1e9, ms.value, by.value


def _scaling(sc):
    """{vel, dcoeff, reac, source} scalings as a double[4]."""
    v = [float(x) for x in sc]
    if len(v) != 4:
        raise ValueError("scaling needs 4 values: vel, dcoeff, reac, source")
    return (C.c_double * 4)(*v)


class FieldSolver:
    """One grid on one GPU (an ``mps_ctx``)."""

    def __init__(self, n, length=(1.0, 1.0, 1.0), p=(1, 1, 1), hierarchy="ref", precision="fp64",
                 flux="upwind", nu1=2, nu2=2, max_cycles=20, tol=1e-11, tol_kind="rms", arith="exact",
                 start="guess", dc_inner=None, omega=None):
        self.lib = M.load()
        self.n = tuple(int(v) for v in (n if not np.isscalar(n) else (n, n, n)))
        self.length = tuple(float(v) for v in (length if not np.isscalar(length) else (length,) * 3))
        g = M.Grid(self.n[0], self.n[1], self.n[2], self.length[0], self.length[1], self.length[2],
                   int(p[0]), int(p[1]), int(p[2]))
        o = M.Opts()
        self.lib.mps_default_opts(C.byref(o))
        o.hierarchy = _NAMES[hierarchy] if isinstance(hierarchy, str) else int(hierarchy)
        o.precision = _PREC[precision] if isinstance(precision, str) else int(precision)
        o.flux = _FLUX[flux] if isinstance(flux, str) else int(flux)
        o.nu1, o.nu2, o.max_cycles = int(nu1), int(nu2), int(max_cycles)
        o.tol = float(tol)
        o.tol_kind = _TOL[tol_kind] if isinstance(tol_kind, str) else int(tol_kind)
        o.arith = {"exact": M.ARITH_EXACT, "fast": M.ARITH_FAST}[arith] if isinstance(arith, str) else int(arith)
        o.start = {"guess": 0, "fmg": 1}[start] if isinstance(start, str) else int(start)
        if dc_inner is not None:
            o.dc_inner = int(dc_inner)
        if omega is not None:
            o.omega = float(omega)
        self.opts = o
        self.max_cycles = int(max_cycles)
        h = C.c_void_p()
        # torch (which owns the device tensors passed in) initialises the HIP
        # runtime before the library does; in the other order torch has been
        # seen to report no device
        try:
            import torch
        except ImportError:
            torch = None
        if torch is not None and torch.cuda.device_count() > 0:
            torch.cuda.init()
        M.check(self.lib.mps_create(C.byref(h), C.byref(g), C.byref(o)))
        self._h = h
        self._keep = []
        self._ext = None        # the context's stream as a torch ExternalStream
        self._torch_call = False
        self._chem = None       # the Chemistry of the last set_chemistry

    # ------------------------------------------------------------ plumbing
    @property
    def shape(self):
        return tuple(v + 2 for v in self.n)

    def close(self):
        if getattr(self, "_h", None):
            self.lib.mps_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def order_torch_in(self):
        """Orders the context's stream after torch's current stream (device-
        side, no host wait): tensors torch is still writing are complete
        before the library reads them. The library itself only orders after
        the null stream."""
        import torch
        if self._ext is None:
            self._ext = torch.cuda.ExternalStream(self.lib.mps_stream(self._h))
        self._ext.wait_stream(torch.cuda.current_stream())
        self._torch_call = True

    def order_torch_out(self):
        """Orders torch's current stream after the context's stream (device-
        side, no host wait): a call on device arrays only may return with its
        work still queued on the context's stream."""
        import torch
        if self._torch_call and self._ext is not None:
            torch.cuda.current_stream().wait_stream(self._ext)
        self._torch_call = False

    def _ptr(self, a, kind, writable=False):
        """Pointer to a caller array. kind: 'g' ghosted, 'i' interior, 'f' face."""
        if a is None:
            return None
        if _is_torch(a):
            import torch
            if not a.is_cuda or a.dtype != torch.float64 or not a.is_contiguous():
                raise ValueError("device arrays must be contiguous float64 tensors on the GPU")
            want = {"g": int(np.prod(self.shape)), "i": int(np.prod(self.n))}.get(kind)
            if want is not None and a.numel() != want:
                raise ValueError("device array has %d elements, expected %d" % (a.numel(), want))
            if not self._torch_call:
                self.order_torch_in()
            return C.c_void_p(a.data_ptr())
        arr = np.asarray(a, dtype=np.float64)
        if kind == "g":
            if arr.shape == self.n and not writable:
                arr = ghosted(arr)
            if arr.shape != self.shape:
                raise ValueError("expected ghosted shape %s, got %s" % (self.shape, arr.shape))
        elif kind == "i" and arr.shape != self.n:
            raise ValueError("expected interior shape %s, got %s" % (self.n, arr.shape))
        if writable:
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.f_contiguous):
                raise ValueError("in/out arrays must be Fortran-ordered float64 numpy arrays (or device tensors)")
            arr = a
        else:
            arr = np.asfortranarray(arr)
        self._keep.append(arr)
        return C.c_void_p(arr.ctypes.data)

    def _release(self):
        self._keep = []
        if self._torch_call:
            self.order_torch_out()

    # ------------------------------------------------------------ set-up
    def set_bc(self, codes, faces=(0.0,) * 6):
        """bc_codes / bcvals of perform_vcycle: codes 'DIRC','ZGRD','FLUX' (others
        act as no-flux), faces left/right (ny,nz), bottom/top (nx,nz), back/front (nx,ny)."""
        nx, ny, nz = self.n
        shapes = [(ny, nz), (ny, nz), (nx, nz), (nx, nz), (nx, ny), (nx, ny)]
        ks = (C.c_int * 6)(*[M.BC_BY_NAME.get(str(c).strip().upper(), M.BC_NONE) if isinstance(c, str) else int(c)
                             for c in codes])
        # the library copies exactly ny*nz, nx*nz or nx*ny values from each
        # device face: a face of another size is refused here, before any call
        for f in range(6):
            v = faces[f]
            if _is_torch(v) and v.numel() != shapes[f][0] * shapes[f][1]:
                raise ValueError("face %d has %d elements, expected %d (%d x %d)"
                                 % (f, v.numel(), shapes[f][0] * shapes[f][1], shapes[f][0], shapes[f][1]))
        ptrs = (C.c_void_p * 6)()
        for f in range(6):
            v = faces[f]
            if _is_torch(v):
                ptrs[f] = self._ptr(v, "f")
            else:
                arr = np.asfortranarray(np.broadcast_to(np.asarray(v, dtype=np.float64), shapes[f]))
                self._keep.append(arr)
                ptrs[f] = arr.ctypes.data
        try:
            M.check(self.lib.mps_set_bc(self._h, ks, ptrs))
        finally:
            self._release()

    def set_coeffs(self, dcoeff=None, velx=None, vely=None, velz=None, reac=None, source=None,
                   timederiv=False, dt=1.0):
        """Operator coefficients (ghosted or interior arrays; None = 0)."""
        try:
            args = [self._ptr(a, "g") for a in (dcoeff, velx, vely, velz, reac, source)]
            M.check(self.lib.mps_set_coeffs(self._h, *args, int(bool(timederiv)), float(dt)))
        finally:
            self._release()

    def set_source(self, source=None):
        """The source of the last set_coeffs replaced (coefficients and hierarchy
        kept): mps_set_source."""
        try:
            M.check(self.lib.mps_set_source(self._h, self._ptr(source, "g")))
        finally:
            self._release()

    def set_case(self, case):
        """Loads a :class:`mps3d_amd.workloads.Case`."""
        self.set_bc(case.bc_codes, case.bc_vals)
        self.set_coeffs(case.dcoeff, case.velx, case.vely, case.velz, case.reac, case.source,
                        case.timederiv, case.dt)

    # ------------------------------------------------------------ compute
    def perform_vcycle(self, soln, mgridsterm=None):
        r = C.c_double()
        try:
            M.check(self.lib.mps_vcycle(self._h, self._ptr(soln, "g", True), self._ptr(mgridsterm, "i"),
                                        C.byref(r)))
        finally:
            self._release()
        return r.value

    def solve(self, soln, raise_on_failure=False):
        """V-cycle loop to the configured tolerance. Returns (ncycles, resnorm, history, status)."""
        n = C.c_int()
        r = C.c_double()
        hist = np.zeros(self.max_cycles)
        try:
            st = self.lib.mps_solve(self._h, self._ptr(soln, "g", True), C.byref(n), C.byref(r),
                                    C.c_void_p(hist.ctypes.data))
        finally:
            self._release()
        M.check(st, allow=(M.OK,) if raise_on_failure else (M.OK, M.NOT_CONVERGED))
        return n.value, r.value, hist[:n.value].copy(), st

    def efield(self, phi, scale=1.0, limiter=False, out=None):
        """update_efield: sets phi's ghosts (in place) and returns E as a
        ghosted (nx+2, ny+2, nz+2, 3) array (host, Fortran order) or into
        `out`, a device tensor of 3 (nz+2, ny+2, nx+2) components."""
        nx, ny, nz = self.n
        E = out if out is not None else np.zeros((nx + 2, ny + 2, nz + 2, 3), order="F")
        try:
            if _is_torch(E):
                if E.numel() != 3 * int(np.prod(self.shape)):
                    raise ValueError("efield needs 3 ghosted components")
                ep = self._ptr(E, None, True)
            else:
                ep = C.c_void_p(E.ctypes.data)
            M.check(self.lib.mps_efield(self._h, self._ptr(phi, "g", True), ep, float(scale), int(bool(limiter))))
        finally:
            self._release()
        return E

    # ------------------------------------------------------------ transport
    def _fields(self, a, ncomp, writable=False):
        """Pointer to ncomp ghosted fields one after the other: a device tensor
        of ncomp * (nx+2)(ny+2)(nz+2) float64 elements, or a Fortran-order
        (nx+2, ny+2, nz+2[, ncomp]) numpy array (in/out arrays must already
        be Fortran-ordered float64)."""
        if a is None:
            return None
        want = ncomp * int(np.prod(self.shape))
        if _is_torch(a):
            import torch
            if not a.is_cuda or a.dtype != torch.float64 or not a.is_contiguous() or a.numel() != want:
                raise ValueError("device arrays must be contiguous float64 GPU tensors of %d elements" % want)
            if not self._torch_call:
                self.order_torch_in()
            return C.c_void_p(a.data_ptr())
        if writable:
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.f_contiguous):
                raise ValueError("in/out arrays must be Fortran-ordered float64 numpy arrays (or device tensors)")
            arr = a
        else:
            arr = np.asfortranarray(a, dtype=np.float64)
        if arr.size != want or arr.shape[:3] != self.shape:
            raise ValueError("expected %d ghosted field(s) of shape %s, got %s" % (ncomp, self.shape, arr.shape))
        self._keep.append(arr)
        return C.c_void_p(arr.ctypes.data)

    def _outputs(self, out, device):
        """vel, dcoeff, reac, source (and prod) outputs: the caller's dict
        entries, or new host arrays / device tensors."""
        nx, ny, nz = self.n
        g = (nx + 2, ny + 2, nz + 2)
        res = dict(out or {})
        if device:
            import torch
            ng = int(np.prod(g))
            dev = device
            for k, m in (("vel", 3), ("dcoeff", 1), ("reac", 1), ("source", 1), ("prod", 1)):
                if k not in res:
                    res[k] = torch.zeros(m * ng, dtype=torch.float64, device=dev)
        else:
            if "vel" not in res:
                res["vel"] = np.zeros(g + (3,), order="F")
            for k in ("dcoeff", "reac", "source", "prod"):
                if k not in res:
                    res[k] = np.zeros(g, order="F")
        return res

    def set_chemistry(self, chem):
        """initializechemistry: the tables of a :class:`Chemistry` to the device."""
        s, keep = chem.struct()
        self._chem = chem
        M.check(self.lib.mps_set_chemistry(self._h, C.byref(s)))
        del keep

    def _need_chem(self):
        if self._chem is None:
            raise M.MpsError(M.ERR_STATE, "set the chemistry first (set_chemistry)")
        return self._chem

    def species_transport(self, specnum, efield, electemp, numden, scaling, soln=None, dt=1.0, timederiv=True,
                          Tg=300.0, Pg=101325.0, out=None):
        """eden_update_transport / ions_update_transport for specnum (1-based):
        returns {"vel", "dcoeff", "reac", "source", "prod"} (ghosted; vel three
        components; prod the unscaled production, interior cells)."""
        dev = efield.device if _is_torch(efield) else None
        o = self._outputs(out, dev)
        ns = self._need_chem().nspecies
        try:
            M.check(self.lib.mps_species_transport(
                self._h, int(specnum), float(dt), int(bool(timederiv)), self._fields(efield, 3),
                self._fields(electemp, 1), self._fields(numden, ns), float(Tg), float(Pg), _scaling(scaling),
                self._fields(soln, 1) if timederiv else None, self._fields(o["vel"], 3, True),
                self._fields(o["dcoeff"], 1, True), self._fields(o["reac"], 1, True),
                self._fields(o["source"], 1, True), self._fields(o["prod"], 1, True)))
        finally:
            self._release()
        return o

    def eenergy_transport(self, efield, electemp, numden, scaling, lscale=1.0, soln=None, dt=1.0, timederiv=True,
                          Tg=300.0, Pg=101325.0, out=None):
        """eenergy_update_transport: {"vel", "dcoeff", "reac", "source"}."""
        dev = efield.device if _is_torch(efield) else None
        o = self._outputs(out, dev)
        o.pop("prod", None)
        ns = self._need_chem().nspecies
        try:
            M.check(self.lib.mps_eenergy_transport(
                self._h, float(dt), int(bool(timederiv)), float(lscale), self._fields(efield, 3),
                self._fields(electemp, 1), self._fields(numden, ns), float(Tg), float(Pg), _scaling(scaling),
                self._fields(soln, 1) if timederiv else None, self._fields(o["vel"], 3, True),
                self._fields(o["dcoeff"], 1, True), self._fields(o["reac"], 1, True),
                self._fields(o["source"], 1, True)))
        finally:
            self._release()
        return o

    def poisson_transport(self, elecden, iondens, scaling, soln=None, dt=1.0, timederiv=True, out=None):
        """poisson_update_transport: {"vel", "dcoeff", "reac", "source"};
        iondens holds the ion species one after the other."""
        dev = elecden.device if _is_torch(elecden) else None
        o = self._outputs(out, dev)
        o.pop("prod", None)
        ch = self._need_chem()
        nions = ch.ionspecmax - ch.ionspecmin + 1
        try:
            M.check(self.lib.mps_poisson_transport(
                self._h, float(dt), int(bool(timederiv)), self._fields(elecden, 1),
                self._fields(iondens, nions) if nions > 0 else None, _scaling(scaling),
                self._fields(soln, 1) if timederiv else None, self._fields(o["vel"], 3, True),
                self._fields(o["dcoeff"], 1, True), self._fields(o["reac"], 1, True),
                self._fields(o["source"], 1, True)))
        finally:
            self._release()
        return o

    def ion_drift_faces(self, specnum, efield, numden, ionsoln, velscale, faces):
        """The FLUX faces of ions_update_bcs for ion specnum (1-based): writes
        the drift outflow into each face of `faces` (six entries: left/right
        (ny,nz), bottom/top (nx,nz), back/front (nx,ny) arrays in Fortran
        order, or device tensors of those sizes; None skips a face)."""
        ns = self._need_chem().nspecies
        nx, ny, nz = self.n
        sizes = (ny * nz, ny * nz, nx * nz, nx * nz, nx * ny, nx * ny)
        ptrs = (C.c_void_p * 6)()
        for k, f in enumerate(faces):
            if f is None:
                continue
            if _is_torch(f):
                if not (f.is_cuda and f.dtype.is_floating_point and f.element_size() == 8 and f.is_contiguous()
                        and f.numel() == sizes[k]):
                    raise ValueError("face %d must be a contiguous float64 GPU tensor of %d elements" % (k, sizes[k]))
                if not self._torch_call:
                    self.order_torch_in()
                ptrs[k] = f.data_ptr()
            else:
                if not (isinstance(f, np.ndarray) and f.dtype == np.float64 and f.flags.f_contiguous
                        and f.size == sizes[k]):
                    raise ValueError("face %d must be a Fortran-ordered float64 array of %d elements" % (k, sizes[k]))
                ptrs[k] = f.ctypes.data
        try:
            M.check(self.lib.mps_ion_drift_faces(self._h, int(specnum), self._fields(efield, 3),
                                                 self._fields(numden, ns), self._fields(ionsoln, 1),
                                                 float(velscale), ptrs))
        finally:
            self._release()
        return faces

    def _codes(self, codes):
        return (C.c_int * 6)(*[M.BC_BY_NAME.get(str(c).strip().upper(), M.BC_NONE) if isinstance(c, str) else int(c)
                               for c in codes])

    def _face_ptrs(self, faces):
        nx, ny, nz = self.n
        want = (ny * nz, ny * nz, nx * nz, nx * nz, nx * ny, nx * ny)
        ptrs = (C.c_void_p * 6)()
        for f in range(6):
            v = faces[f]
            if v is None:
                continue
            if not _is_torch(v) or v.numel() != want[f] or not v.is_contiguous():
                raise ValueError("face %d must be a contiguous device tensor of %d values" % (f, want[f]))
            ptrs[f] = self._ptr(v, "f", True)
        return ptrs

    def eden_faces(self, codes, elecden, electemp, velscale, faces):
        """eden_update_bcs (plasma_pde/edeneq.f90:203-358): the FLUX faces of
        `faces` (six device tensors or None, rewritten in place) get the thermal
        outflow of the scaled electron density: mps_eden_faces."""
        try:
            M.check(self.lib.mps_eden_faces(self._h, self._codes(codes), self._fields(elecden, 1),
                                            self._fields(electemp, 1), float(velscale), self._face_ptrs(faces)))
        finally:
            self._release()

    def neutral_faces(self, codes, neutralden, molmass, Tg, velscale, faces):
        """neutrals_update_bcs (plasma_pde/neutralseq.f90:206-358): mps_neutral_faces."""
        try:
            M.check(self.lib.mps_neutral_faces(self._h, self._codes(codes), self._fields(neutralden, 1),
                                               float(molmass), float(Tg), float(velscale), self._face_ptrs(faces)))
        finally:
            self._release()

    def eenergy_faces(self, codes, elecenergy, elecden, eenergyscale, bcparams, velscale, faces):
        """eenergy_update_bcs (plasma_pde/eenergyeq.f90:208-392): mps_eenergy_faces."""
        par = (C.c_double * 6)(*[float(v) for v in bcparams])
        try:
            M.check(self.lib.mps_eenergy_faces(self._h, self._codes(codes), self._fields(elecenergy, 1),
                                               self._fields(elecden, 1), float(eenergyscale), par, float(velscale),
                                               self._face_ptrs(faces)))
        finally:
            self._release()

    def floor(self, soln, vmin):
        """The end of eden_solve (plasma_pde/edeneq.f90:509-523): interior values
        below vmin raised to vmin, then the ghost values: mps_floor."""
        try:
            M.check(self.lib.mps_floor(self._h, self._ptr(soln, "g", True), float(vmin)))
        finally:
            self._release()

    def electron_temp(self, eenergysoln, elecden, eenergyscale, etempmin, electemp):
        """update_electrontemp: writes electemp's interior and clamps
        eenergysoln (both in/out)."""
        try:
            M.check(self.lib.mps_electron_temp(self._h, self._fields(eenergysoln, 1, True),
                                               self._fields(elecden, 1), float(eenergyscale), float(etempmin),
                                               self._fields(electemp, 1, True)))
        finally:
            self._release()
        return electemp

    def solve_rhs(self, b, soln, raise_on_failure=False):
        """A x = b for a right-hand side b (interior) given directly, with the
        coefficients of the last set_coeffs (their source and the faces are
        not used). Returns (ncycles, resnorm, history, status)."""
        n = C.c_int()
        r = C.c_double()
        hist = np.zeros(self.max_cycles)
        try:
            st = self.lib.mps_solve_rhs(self._h, self._ptr(b, "i"), self._ptr(soln, "g", True), C.byref(n),
                                        C.byref(r), C.c_void_p(hist.ctypes.data))
        finally:
            self._release()
        M.check(st, allow=(M.OK,) if raise_on_failure else (M.OK, M.NOT_CONVERGED))
        return n.value, r.value, hist[:n.value].copy(), st

    def last_timing(self):
        ms = C.c_double()
        n = C.c_int()
        M.check(self.lib.mps_last_timing(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def last_vcycles(self):
        """V-cycles of the last solve (FP32 ones under fp32dc; a full-multigrid
        start counts as one)."""
        n = C.c_int()
        M.check(self.lib.mps_last_vcycles(self._h, C.byref(n)))
        return n.value

    def time_kernel(self, kernel=0, fp32=False, reps=20):
        """Mean launch time (ms), algorithmic bytes per launch and symbol name of
        one hot kernel (0 colour pass, 1 residual, 2 restriction, 3 prolongation)."""
        ms, by = C.c_double(), C.c_double()
        name = C.create_string_buffer(64)
        M.check(self.lib.mps_time_kernel(self._h, int(kernel), int(bool(fp32)), int(reps), C.byref(ms),
                                         C.byref(by), name, 64))
        return ms.value, by.value, name.value.decode()

    def time_kernel_in_solve(self, soln, kernel=0, fp32=False):
        """Runs solve(soln) with every finest-level launch of one hot kernel
        (0 sweep, 1 residual or FP64 defect, 2 restriction, 3 prolongation)
        timed by its own HIP events; returns (mean ms per launch, launches,
        V-cycles)."""
        ms, nl, nc = C.c_double(), C.c_int(), C.c_int()
        try:
            st = self.lib.mps_time_in_solve(self._h, int(kernel), int(bool(fp32)), self._ptr(soln, "g", True),
                                            C.byref(nc), C.byref(ms), C.byref(nl))
        finally:
            self._release()
        M.check(st, allow=(M.OK, M.NOT_CONVERGED))
        return ms.value, nl.value, nc.value

    def find_AX(self, x):
        ax, dg, of = (np.zeros(self.n, order="F") for _ in range(3))
        try:
            M.check(self.lib.mps_apply(self._h, self._ptr(x, "g"), self._ptr(ax, "i", True),
                                       self._ptr(dg, "i", True), self._ptr(of, "i", True)))
        finally:
            self._release()
        return ax, dg, of

    def apply_transpose(self, x):
        out = np.zeros(self.n, order="F")
        try:
            M.check(self.lib.mps_apply_transpose(self._h, self._ptr(x, "g"), self._ptr(out, "i", True)))
        finally:
            self._release()
        return out

    def find_bvec(self, mgridsterm=None, applybc=True):
        b = np.zeros(self.n, order="F")
        try:
            M.check(self.lib.mps_bvec(self._h, self._ptr(mgridsterm, "i"), int(bool(applybc)),
                                      self._ptr(b, "i", True)))
        finally:
            self._release()
        return b

    def restrict(self, fine, out=None):
        """restriction (mgridlib/mgridsteps.f90:11-53): the 8-child mean of an
        interior (nx, ny, nz) array onto (nx/2, ny/2, nz/2). A numpy input
        returns a numpy array; with device tensors, pass `out` (nz/2, ny/2,
        nx/2)."""
        nc = tuple(v // 2 for v in self.n)
        if _is_torch(fine):
            if out is None or out.numel() != int(np.prod(nc)):
                raise ValueError("device input needs an output tensor of %d elements" % int(np.prod(nc)))
            dst, ret = out, out
        else:
            dst = ret = np.zeros(nc, order="F")
        try:
            p_out = self._ptr(dst, "c", True) if _is_torch(dst) else C.c_void_p(dst.ctypes.data)
            M.check(self.lib.mps_restrict(self._h, self._ptr(fine, "i"), p_out))
        finally:
            self._release()
        return ret

    def prolong1(self, coarse):
        """prolongation1 (mgridlib/mgridsteps.f90:170-216): the interior
        (nx, ny, nz) interpolant of a ghosted (nx/2+2, ny/2+2, nz/2+2) coarse
        array, whose ghost layer is read as given."""
        nc = tuple(v // 2 + 2 for v in self.n)
        cg = np.asfortranarray(np.asarray(coarse, dtype=np.float64))
        if cg.shape != nc:
            raise ValueError("expected coarse ghosted shape %s, got %s" % (nc, cg.shape))
        fine = np.zeros(self.n, order="F")
        M.check(self.lib.mps_prolong1(self._h, C.c_void_p(cg.ctypes.data), C.c_void_p(fine.ctypes.data)))
        return fine

    def smooth(self, soln, niter=2, mgridsterm=None):
        res = np.zeros(self.n, order="F")
        try:
            M.check(self.lib.mps_smooth(self._h, self._ptr(soln, "g", True), self._ptr(mgridsterm, "i"),
                                        int(niter), self._ptr(res, "i", True)))
        finally:
            self._release()
        return res

    def fill_ghosts(self, soln):
        try:
            M.check(self.lib.mps_fill_ghosts(self._h, self._ptr(soln, "g", True)))
        finally:
            self._release()

    def gmres(self, b, x, m=30, restarts=20, tol=1e-10, precond=1, raise_on_failure=False):
        """precond 0: the reference's performgmres (left, one RB sweep, small grids);
        precond 1: FGMRES(m) right-preconditioned by one V-cycle (the FP32 cycle under
        precision "fp32dc"). Returns (iters, relres, status)."""
        it = C.c_int()
        rr = C.c_double()
        try:
            st = self.lib.mps_gmres(self._h, self._ptr(b, "i"), self._ptr(x, "g", True), int(m), int(restarts),
                                    float(tol), int(precond), C.byref(it), C.byref(rr))
        finally:
            self._release()
        M.check(st, allow=(M.OK,) if raise_on_failure else (M.OK, M.NOT_CONVERGED))
        return it.value, rr.value, st

    def solve_transpose(self, g, lam):
        n = C.c_int()
        r = C.c_double()
        try:
            st = self.lib.mps_solve_transpose(self._h, self._ptr(g, "i"), self._ptr(lam, "g", True), C.byref(n),
                                              C.byref(r))
        finally:
            self._release()
        M.check(st, allow=(M.OK, M.NOT_CONVERGED))
        return n.value, r.value, st
