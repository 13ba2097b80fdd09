"""mrcpp_amd: MRCPP's convolution-operator application on the AMD MI355X.

Python mirror of the MRCPP interface for this path (names, argument order and
meaning as in MRCPP's C++ API):

    mra = MultiResolutionAnalysis((-32, 32), order=7)     # BoundingBox + LegendreBasis
    f = FunctionTree(mra)
    g = GaussExp(); g.append(GaussFunc(beta, coef, pos))
    build_grid(f, g); project(prec, f, g)
    P = PoissonOperator(mra, prec)        # or HelmholtzOperator(mra, mu, prec)
    out = FunctionTree(mra); apply(prec, out, P, f)
    dot(out, f), out.evalf(r), out.integrate(), out.getSquareNorm()

Everything calls the C ABI of include/mrcpp_amd.h in
mrcpp_amd/lib/libmrcpp_amd.so; the band sums run as gfx950 kernels.  There is
no CPU fallback for apply: without the library or a GPU it raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmrcpp_amd.so")
FILTER_DIR = os.path.join(HERE, "share", "mwfilters")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_lib = None

# name: (restype, argtypes) of every symbol declared in include/mrcpp_amd.h
SIGNATURES = {
    "mrcpp_last_error": (ctypes.c_char_p, []),
    "mrcpp_set_filter_dir": (None, [ctypes.c_char_p]),
    "mrcpp_mra_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "mrcpp_mra_create_box": (ctypes.c_int, [ctypes.c_int, _ip, _ip, _dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "mrcpp_mra_destroy": (None, [_vp]),
    "mrcpp_mra_info": (ctypes.c_int, [_vp, _ip, _dp, _ip]),
    "mrcpp_tree_create": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "mrcpp_tree_destroy": (None, [_vp]),
    "mrcpp_tree_build_grid_gaussians": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp]),
    "mrcpp_tree_project_gaussians": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp]),
    "mrcpp_tree_project_gaussians_ex": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int]),
    "mrcpp_tree_num_nodes": (ctypes.c_longlong, [_vp]),
    "mrcpp_tree_square_norm": (ctypes.c_double, [_vp]),
    "mrcpp_tree_export": (ctypes.c_int, [_vp, _ip, _dp, _dp]),
    "mrcpp_tree_evalf": (ctypes.c_double, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "mrcpp_tree_integrate": (ctypes.c_double, [_vp]),
    "mrcpp_tree_dot": (ctypes.c_double, [_vp, _vp]),
    "mrcpp_add": (ctypes.c_int, [ctypes.c_double, _vp, ctypes.c_int, _dp, ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int]),
    "mrcpp_multiply": (ctypes.c_int, [ctypes.c_double, _vp, ctypes.c_int, _dp, ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int]),
    "mrcpp_multiply_ex": (ctypes.c_int, [ctypes.c_double, _vp, ctypes.c_int, _dp, ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "mrcpp_tree_rescale": (ctypes.c_int, [_vp, ctypes.c_double]),
    "mrcpp_apply_components": (ctypes.c_int, [ctypes.c_double, ctypes.POINTER(_vp), _vp, ctypes.POINTER(_vp), ctypes.c_int, _dp, ctypes.c_int,
                                              ctypes.c_int]),
    "mrcpp_tree_save": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "mrcpp_tree_load": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "mrcpp_tree_save_txt": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "mrcpp_tree_load_txt": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "mrcpp_operator_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.POINTER(_vp)]),
    "mrcpp_operator_destroy": (None, [_vp]),
    "mrcpp_operators_create": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.POINTER(_vp)]),
    "mrcpp_operator_build_phases": (ctypes.c_int, [_vp, _dp]),
    "mrcpp_operator_size": (ctypes.c_int, [_vp]),
    "mrcpp_operator_build_seconds": (ctypes.c_double, [_vp]),
    "mrcpp_operator_kernel": (ctypes.c_int, [_vp, _dp, _dp]),
    "mrcpp_operator_bandwidths": (ctypes.c_int, [_vp, ctypes.c_double, _ip, ctypes.c_int]),
    "mrcpp_operator_node": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "mrcpp_kernel_terms": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          _dp, _dp, ctypes.c_int]),
    "mrcpp_apply": (ctypes.c_int, [ctypes.c_double, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _dp]),
    "mrcpp_apply_prec": (ctypes.c_int, [ctypes.c_double, _vp, _vp, _vp, ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    "mrcpp_set_projection_target": (ctypes.c_int, [ctypes.c_int]),
    "mrcpp_set_operator_target": (ctypes.c_int, [ctypes.c_int]),
    "mrcpp_operator_built_on_device": (ctypes.c_int, [_vp]),
    "mrcpp_last_projection_on_device": (ctypes.c_int, []),
    "mrcpp_apply_batch": (ctypes.c_int, [ctypes.c_int, _dp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _dp]),
    "mrcpp_apply_batch_predicted": (ctypes.c_int, [ctypes.c_int, _dp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp,
                                                   _dp, ctypes.POINTER(ctypes.c_int), ctypes.c_int, _dp]),
    "mrcpp_predictor_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "mrcpp_predictor_destroy": (ctypes.c_int, [_vp]),
    "mrcpp_predictor_num_params": (ctypes.c_int, []),
    "mrcpp_predictor_set_params": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int]),
    "mrcpp_predictor_get_params": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int]),
    "mrcpp_predictor_set_threshold": (ctypes.c_int, [_vp, ctypes.c_double]),
    "mrcpp_predictor_predict": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
    "mrcpp_predictor_train": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_float)]),
    "mrcpp_samples_count": (ctypes.c_int, []),
    "mrcpp_samples_fetch": (ctypes.c_int, [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_int]),
    "mrcpp_set_sweep_precision": (ctypes.c_int, [ctypes.c_int]),
    "mrcpp_get_sweep_precision": (ctypes.c_int, []),
    "mrcpp_set_level_ahead": (ctypes.c_int, [ctypes.c_int]),
    "mrcpp_level_ahead": (ctypes.c_int, []),
    "mrcpp_last_apply_levels": (ctypes.c_int, [_ip, ctypes.c_int]),
    # device layer (include/mwg.h) used by the harness
    "mwg_last_error": (ctypes.c_char_p, []),
    "mwg_device_count": (ctypes.c_int, []),
    "mwg_synchronize": (ctypes.c_int, []),
    "mwg_profile": (ctypes.c_int, [ctypes.c_int]),
    "mwg_profile_kinds": (ctypes.c_int, [ctypes.c_uint]),
    "mwg_set_band_layout": (ctypes.c_int, [ctypes.c_int]),
    "mwg_band_layout": (ctypes.c_int, []),
    "mwg_set_band_launch": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "mwg_band_launch": (None, [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "mrcpp_autotune_band": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int]),
    "mwg_profile_read": (ctypes.c_int, [_dp, ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]),
    "mwg_release_pool": (ctypes.c_int, []),
    "mwg_init": (ctypes.c_int, [ctypes.c_int]),
    "mwg_malloc": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_size_t]),
    "mwg_free": (ctypes.c_int, [_vp]),
    "mwg_upload": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "mwg_download": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "mwg_mlp_forward": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp]),
    "mwg_mlp_gradient": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_float, _vp]),
    "mwg_mlp_adam": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
}

KERNEL_KINDS = ("screen", "generate", "band", "reduce", "reconstruct", "compress", "norms", "groups", "project", "arith")


class MRCPPError(RuntimeError):
    pass


def lib():
    """Load libmrcpp_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MRCPPError("libmrcpp_amd.so is missing: run python -m mrcpp_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L.mrcpp_set_filter_dir(FILTER_DIR.encode())
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise MRCPPError(lib().mrcpp_last_error().decode())


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


class GaussFunc:
    """coef * exp(-alpha |r - pos|^2) (src/functions/GaussFunc.h)."""

    def __init__(self, alpha, coef, pos=(0.0, 0.0, 0.0)):
        self.alpha = float(alpha)
        self.coef = float(coef)
        self.pos = tuple(float(x) for x in pos)


class GaussExp:
    """Sum of GaussFunc terms (src/functions/GaussExp.h)."""

    def __init__(self, funcs=()):
        self.funcs = list(funcs)

    def append(self, g):
        self.funcs.append(g)

    def size(self):
        return len(self.funcs)

    def arrays(self):
        c = np.array([g.coef for g in self.funcs], dtype=np.float64)
        a = np.array([g.alpha for g in self.funcs], dtype=np.float64)
        r = np.array([g.pos for g in self.funcs], dtype=np.float64).reshape(-1, 3)
        return c, a, np.ascontiguousarray(r)

    @classmethod
    def from_sources(cls, src):
        """From a workloads.GaussianSources set."""
        return cls(GaussFunc(b, c, p) for c, b, p in zip(src.coefs, src.exponents, src.centers))


class MultiResolutionAnalysis:
    """MultiResolutionAnalysis<3>(BoundingBox<3>({lo, hi}), LegendreBasis(order), depth), or with
    box = (scale, corner, nboxes[, sf]) the world BoundingBox<3>(scale, corner, nboxes, sf):
    nboxes[d] root boxes from translation corner[d] at root scale `scale`."""

    def __init__(self, box, order, depth=30):
        self.order = int(order)
        h = _vp()
        if len(box) == 2:
            self.box = (int(box[0]), int(box[1]))
            _check(lib().mrcpp_mra_create(self.box[0], self.box[1], self.order, int(depth), ctypes.byref(h)))
        else:
            scale, corner, nb = int(box[0]), np.asarray(box[1], dtype=np.int32), np.asarray(box[2], dtype=np.int32)
            sf = np.asarray(box[3] if len(box) > 3 else (1.0, 1.0, 1.0), dtype=np.float64)
            self.box = (scale, tuple(corner.tolist()), tuple(nb.tolist()), tuple(sf.tolist()))
            _check(lib().mrcpp_mra_create_box(scale, _p(corner, _ip), _p(nb, _ip), _p(sf), self.order, int(depth), ctypes.byref(h)))
        self._h = h
        n0, sf, ms = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
        _check(lib().mrcpp_mra_info(h, ctypes.byref(n0), ctypes.byref(sf), ctypes.byref(ms)))
        self.root_scale, self.scaling_factor, self.max_scale = n0.value, sf.value, ms.value

    def getOrder(self):
        return self.order

    def getMaxScale(self):
        return self.max_scale

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mrcpp_mra_destroy(self._h)
            self._h = None


class FunctionTree:
    """FunctionTree<3>(mra)."""

    def __init__(self, mra, name="nn"):
        self.mra = mra
        self.name = name
        h = _vp()
        _check(lib().mrcpp_tree_create(mra._h, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mrcpp_tree_destroy(self._h)
            self._h = None

    def getNNodes(self):
        return int(lib().mrcpp_tree_num_nodes(self._h))

    def getSquareNorm(self):
        return float(lib().mrcpp_tree_square_norm(self._h))

    def evalf(self, r):
        return float(lib().mrcpp_tree_evalf(self._h, float(r[0]), float(r[1]), float(r[2])))

    def integrate(self):
        return float(lib().mrcpp_tree_integrate(self._h))

    def rescale(self, c):
        """FunctionTree::rescale: end nodes times c, bottom-up transform, norms."""
        _check(lib().mrcpp_tree_rescale(self._h, float(c)))

    def saveTree(self, file):
        """Binary file "<file>.tree" (FunctionTree::saveTree)."""
        _check(lib().mrcpp_tree_save(self._h, os.fsencode(file)))

    def loadTree(self, file):
        """Read "<file>.tree"; the tree must have the MRA it was saved with."""
        _check(lib().mrcpp_tree_load(self._h, os.fsencode(file)))

    def saveTreeTXT(self, file):
        """MADNESS-convention text file (FunctionTree::saveTreeTXT)."""
        _check(lib().mrcpp_tree_save_txt(self._h, os.fsencode(file)))

    def loadTreeTXT(self, file):
        """Read a MADNESS-convention text file (FunctionTree::loadTreeTXT)."""
        _check(lib().mrcpp_tree_load_txt(self._h, os.fsencode(file)))

    def export(self, with_coefs=True):
        """(idx[N,5] = scale, lx, ly, lz, is_end; coefs[N, 8 (k+1)^3] or None; norms[N, 9])."""
        n = self.getNNodes()
        kp1_d = (self.mra.order + 1) ** 3
        idx = np.zeros((n, 5), dtype=np.int32)
        coefs = np.zeros((n, 8 * kp1_d)) if with_coefs else None
        norms = np.zeros((n, 9))
        got = lib().mrcpp_tree_export(self._h, _p(idx, _ip), _p(coefs) if with_coefs else None, _p(norms))
        if got < 0:
            raise MRCPPError(lib().mrcpp_last_error().decode())
        return idx, coefs, norms


class ConvolutionOperator:
    POISSON, HELMHOLTZ = 0, 1

    def __init__(self, mra, kind, mu, prec):
        self.mra = mra
        h = _vp()
        _check(lib().mrcpp_operator_create(mra._h, kind, float(mu), float(prec), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mrcpp_operator_destroy(self._h)
            self._h = None

    def size(self):
        return int(lib().mrcpp_operator_size(self._h))

    def build_seconds(self):
        return float(lib().mrcpp_operator_build_seconds(self._h))

    def built_on_device(self):
        """Whether the 2D operator trees were formed on the GPU."""
        return bool(lib().mrcpp_operator_built_on_device(self._h))

    def build_phases(self):
        """Seconds of the construction phases: kernel trees and term set-up;
        per refinement level the packing of the GPU's inputs, the GPU round
        trips and the host's share (a host build counts its whole 2D tree
        construction here); bottom-up transforms and translation caches."""
        out = np.zeros(5)
        _check(lib().mrcpp_operator_build_phases(self._h, _p(out)))
        return dict(zip(("kernel_trees", "pack", "device", "host_levels", "finish"), map(float, out)))

    def kernel(self):
        m = self.size()
        a, c = np.zeros(m), np.zeros(m)
        _check(lib().mrcpp_operator_kernel(self._h, _p(a), _p(c)))
        return a, c

    def bandwidths(self, prec, max_depth=32):
        out = np.zeros((self.size(), max_depth, 5), dtype=np.int32)
        if lib().mrcpp_operator_bandwidths(self._h, float(prec), _p(out, _ip), max_depth) < 0:
            raise MRCPPError(lib().mrcpp_last_error().decode())
        return out

    def node(self, term, depth, l, comp):
        kp1 = self.mra.order + 1
        mat = np.zeros(kp1 * kp1)
        nrm = ctypes.c_double()
        if lib().mrcpp_operator_node(self._h, term, depth, l, comp, _p(mat), ctypes.byref(nrm)) != 0:
            return None, None
        return mat.reshape(kp1, kp1, order="F"), nrm.value


class PoissonOperator(ConvolutionOperator):
    """PoissonOperator(mra, prec)."""

    def __init__(self, mra, prec):
        super().__init__(mra, self.POISSON, 0.0, prec)


class HelmholtzOperator(ConvolutionOperator):
    """HelmholtzOperator(mra, mu, prec)."""

    def __init__(self, mra, mu, prec):
        super().__init__(mra, self.HELMHOLTZ, mu, prec)


def build_operators(mra, specs, target="device"):
    """Several operators on one MRA built together: specs is a list of
    (mu, prec), mu None or <= 0 for Poisson, else Helmholtz(mu).  With target
    "device" the 2D trees of all their terms are formed in one GPU launch per
    refinement level; "host" builds them one by one on the host cores.  Each
    operator equals its own constructor's, bit for bit."""
    n = len(specs)
    mus = np.array([0.0 if mu is None else float(mu) for mu, _ in specs], dtype=np.float64)
    precs = np.array([float(p) for _, p in specs], dtype=np.float64)
    hs = (_vp * max(n, 1))()
    _check(lib().mrcpp_operators_create(mra._h, n, _p(mus), _p(precs), _PROJ_TARGETS[target], hs))
    out = []
    for i in range(n):
        mu = float(mus[i])
        op = ConvolutionOperator.__new__(PoissonOperator if mu <= 0.0 else HelmholtzOperator)
        op.mra = mra
        op._h = _vp(hs[i])
        out.append(op)
    return out


def kernel_terms(kind, mu, eps, r_min, r_max):
    a, c = np.zeros(2000), np.zeros(2000)
    n = lib().mrcpp_kernel_terms(int(kind), float(mu), float(eps), float(r_min), float(r_max), _p(a), _p(c), 2000)
    if n < 0:
        raise MRCPPError(lib().mrcpp_last_error().decode())
    return a[:n].copy(), c[:n].copy()


def build_grid(out, gexp):
    c, a, r = gexp.arrays()
    _check(lib().mrcpp_tree_build_grid_gaussians(out._h, len(c), _p(c), _p(a), _p(r)))


def project(prec, out, gexp, maxIter=-1, absPrec=False):
    """project(prec, tree, GaussExp, maxIter, absPrec): on the GPU when one is
    present (set_projection_target), else on the host cores."""
    c, a, r = gexp.arrays()
    _check(lib().mrcpp_tree_project_gaussians_ex(out._h, float(prec), len(c), _p(c), _p(a), _p(r), int(maxIter), int(bool(absPrec))))


_PROJ_TARGETS = {"auto": 0, "host": 1, "device": 2}


def set_projection_target(target):
    """Where later project() calls run: "auto" (the GPU when present), "host" or "device"."""
    _check(lib().mrcpp_set_projection_target(_PROJ_TARGETS[target]))


def set_operator_target(target):
    """Where later operator constructions form their 2D operator trees: "auto"
    (the host cores), "host" or "device" (bit-identical results)."""
    _check(lib().mrcpp_set_operator_target(_PROJ_TARGETS[target]))


def last_projection_on_device():
    """Whether the last project() ran on the GPU."""
    return bool(lib().mrcpp_last_projection_on_device())


def _terms(args, name):
    """(c_i, tree_i) terms and the trailing (maxIter, absPrec) of add/multiply's
    two forms: (a, tree_a, b, tree_b, ...) or ([(c, tree), ...], ...)."""
    if args and isinstance(args[0], (list, tuple)):
        return [(float(c), t) for c, t in args[0]], args[1:]
    if len(args) < 4:
        raise TypeError("%s(prec, out, a, tree_a, b, tree_b[, maxIter, absPrec]) or %s(prec, out, [(c, tree), ...])" % (name, name))
    return [(float(args[0]), args[1]), (float(args[2]), args[3])], args[4:]


def _arith(fn, prec, out, terms, max_iter, abs_prec):
    c = np.ascontiguousarray([t[0] for t in terms], dtype=np.float64)
    hs = (_vp * max(len(terms), 1))(*[t[1]._h for t in terms])
    _check(fn(float(prec), out._h, len(terms), _p(c), hs, int(max_iter), int(bool(abs_prec))))


def add(prec, out, *args, maxIter=-1, absPrec=False):
    """add(prec, out, a, tree_a, b, tree_b[, maxIter, absPrec]) or
    add(prec, out, [(c, tree), ...][, maxIter, absPrec]) on the GPU."""
    terms, rest = _terms(args, "add")
    if rest:
        maxIter = rest[0]
    if len(rest) > 1:
        absPrec = rest[1]
    _arith(lib().mrcpp_add, prec, out, terms, maxIter, absPrec)


def multiply(prec, out, *args, maxIter=-1, absPrec=False, useMaxNorms=False):
    """multiply(prec, out, c, tree_a, tree_b[, maxIter, absPrec, useMaxNorms]) or
    multiply(prec, out, [(c, tree), ...][, maxIter, absPrec, useMaxNorms]) on the GPU;
    useMaxNorms refines with MRCPP's MultiplicationAdaptor (two trees)."""
    if args and not isinstance(args[0], (list, tuple)):
        if len(args) < 3:
            raise TypeError("multiply(prec, out, c, tree_a, tree_b[, maxIter, absPrec, useMaxNorms])")
        args = ([(args[0], args[1]), (1.0, args[2])],) + tuple(args[3:])
    terms, rest = _terms(args, "multiply")
    if rest:
        maxIter = rest[0]
    if len(rest) > 1:
        absPrec = rest[1]
    if len(rest) > 2:
        useMaxNorms = rest[2]
    c = np.ascontiguousarray([t[0] for t in terms], dtype=np.float64)
    hs = (_vp * max(len(terms), 1))(*[t[1]._h for t in terms])
    _check(lib().mrcpp_multiply_ex(float(prec), out._h, len(terms), _p(c), hs, int(maxIter), int(bool(absPrec)), int(bool(useMaxNorms))))


def dot(bra, ket):
    return float(lib().mrcpp_tree_dot(bra._h, ket._h))


STAT_KEYS = ("n_applied", "n_pairs", "n_generated", "n_output_nodes", "seconds", "n_levels", "t_screen", "t_generate",
             "t_band", "t_setup", "t_transform", "t_host", "t_cleanup", "n_ahead")


def device_synchronize():
    if lib().mwg_synchronize() != 0:
        raise MRCPPError(lib().mwg_last_error().decode())


def profile(enable, kinds=None):
    """Start (True) or stop recording per-kernel HIP-event timings; kinds
    (names from KERNEL_KINDS) limits the events to those kernels."""
    if kinds is None:
        rc = lib().mwg_profile(int(bool(enable)))
    else:
        mask = 0
        for k in kinds:
            mask |= 1 << KERNEL_KINDS.index(k)
        rc = lib().mwg_profile_kinds(mask if enable else 0)
    if rc != 0:
        raise MRCPPError(lib().mwg_last_error().decode())


def profile_read():
    """{kind: (total ms, launches)} of the kernels launched since profile(True)."""
    n = len(KERNEL_KINDS)
    ms = np.zeros(n)
    cnt = np.zeros(n, dtype=np.int64)
    if lib().mwg_profile_read(_p(ms), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), n) != 0:
        raise MRCPPError(lib().mwg_last_error().decode())
    return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_KINDS)}


SWEEP_PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2, "fp64_vector": 3}


def set_sweep_precision(name):
    """Sweep precision of later apply() calls on this thread: 'fp64' (default,
    FP64 matrix cores), 'fp32', 'bf16x3' (band sum accumulated in FP64 in
    both) or 'fp64_vector' (FP64 sweeps on the vector ALUs)."""
    if name not in SWEEP_PRECISIONS:
        raise ValueError("sweep precision must be one of %s" % sorted(SWEEP_PRECISIONS))
    _check(lib().mrcpp_set_sweep_precision(SWEEP_PRECISIONS[name]))


def sweep_precision():
    code = lib().mrcpp_get_sweep_precision()
    return {v: k for k, v in SWEEP_PRECISIONS.items()}[code]


def set_level_ahead(on):
    """Whether later apply() calls on this thread form each breadth-first
    step's next level and its pair screen on the GPU before the host's split
    (single plain trees, steps of at most 2048 nodes; taken only where the
    host's split agrees bit for bit).  Default False; the same trees either way."""
    _check(lib().mrcpp_set_level_ahead(1 if on else 0))


def level_ahead():
    return bool(lib().mrcpp_level_ahead())


BAND_LAYOUTS = {"auto": 0, "tiles": 1, "rows": 2}


def set_band_layout(name):
    """Work distribution of the order-7 FP64/FP32 band kernel (process-wide):
    'rows' (each wave owns 16 rows of every sweep, one workgroup barrier per
    group), 'tiles' (sweep tiles shared out over the waves, two barriers per
    group) or 'auto' (the default: 'rows').  Results are bit-identical."""
    if name not in BAND_LAYOUTS:
        raise ValueError("band layout must be one of %s" % sorted(BAND_LAYOUTS))
    if lib().mwg_set_band_layout(BAND_LAYOUTS[name]) != 0:
        raise MRCPPError(lib().mwg_last_error().decode())


def band_layout():
    return {v: k for k, v in BAND_LAYOUTS.items()}[lib().mwg_band_layout()]


def set_band_launch(unit_groups=32, workgroups_per_cu=0):
    """Minimum groups per work unit of the band sum (8..1024, default 32) and
    workgroups per CU of the persistent band kernels (0: as many as fit).
    Process-wide; results change only in the last bits (the order of a
    node's partial sums)."""
    if lib().mwg_set_band_launch(int(unit_groups), int(workgroups_per_cu)) != 0:
        raise MRCPPError(lib().mwg_last_error().decode())


def band_launch():
    ug, wg = ctypes.c_int(), ctypes.c_int()
    lib().mwg_band_launch(ctypes.byref(ug), ctypes.byref(wg))
    return ug.value, wg.value


_RUN_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p)


def autotune_band(run, repeats=3):
    """Time run() (the caller's workload, e.g. one apply) under each launch
    shape of the band sum -- layout x minimum unit groups (24, 32, 48) x
    workgroups per CU (as many as fit, or 3) -- and leave the fastest set.
    Returns {"best": {...}, "candidates": [...]} with seconds per run."""
    errors = []

    def call(_ctx):
        if errors:
            return
        try:
            run()
        except BaseException as e:  # re-raised after the tuner returns
            errors.append(e)

    cb = _RUN_CB(call)
    best = np.zeros(4)
    table = np.zeros(4 * 64)
    before = (band_layout(), band_launch())
    n = lib().mrcpp_autotune_band(ctypes.cast(cb, ctypes.c_void_p), None, int(repeats), _p(best), _p(table), 64)
    if errors:  # the workload failed: no timing is valid, the settings go back
        set_band_layout(before[0])
        set_band_launch(*before[1])
        raise errors[0]
    if n < 0:
        raise MRCPPError(lib().mrcpp_last_error().decode())
    names = {v: k for k, v in BAND_LAYOUTS.items()}

    def row(v):
        return {"layout": names[int(v[0])], "unit_groups": int(v[1]), "workgroups_per_cu": int(v[2]), "seconds": float(v[3])}

    return {"best": row(best), "candidates": [row(table[4 * i:4 * i + 4]) for i in range(min(n, 64))]}


_fp = ctypes.POINTER(ctypes.c_float)
PREDICTOR_FEATURES = 12


class RefinePredictor:
    """12-64-64-1 perceptron (FP32, on the device) predicting which output
    nodes the adaptive apply refines; apply(..., predictor=P) grows the output
    grid with it after the second step, the wavelet test staying the judge."""

    def __init__(self, params=None, threshold=None):
        h = _vp()
        _check(lib().mrcpp_predictor_create(ctypes.byref(h)))
        self._h = h
        if params is not None:
            self.params = params
        if threshold is not None:
            self.threshold = threshold

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mrcpp_predictor_destroy(self._h)
            self._h = None

    @property
    def params(self):
        p = np.zeros(lib().mrcpp_predictor_num_params(), dtype=np.float32)
        _check(lib().mrcpp_predictor_get_params(self._h, _p(p, _fp), len(p)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().mrcpp_predictor_set_params(self._h, _p(p, _fp), len(p)))

    @property
    def threshold(self):
        return getattr(self, "_thr", 0.5)

    @threshold.setter
    def threshold(self, t):
        _check(lib().mrcpp_predictor_set_threshold(self._h, float(t)))
        self._thr = float(t)

    def predict(self, features):
        x = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, PREDICTOR_FEATURES)
        out = np.zeros(len(x), dtype=np.float32)
        _check(lib().mrcpp_predictor_predict(self._h, _p(x, _fp), len(x), _p(out, _fp)))
        return out

    def train(self, features, labels, epochs=20, batch=4096, lr=2e-3, pos_weight=3.0, seed=1):
        """Adam on the device; returns the loss after every epoch."""
        x = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, PREDICTOR_FEATURES)
        y = np.ascontiguousarray(labels, dtype=np.float32)
        losses = np.zeros(epochs, dtype=np.float32)
        _check(lib().mrcpp_predictor_train(self._h, _p(x, _fp), _p(y, _fp), len(y), int(epochs), int(batch), float(lr), float(pos_weight),
                                           int(seed), _p(losses, _fp)))
        return losses


def _batch_args(tasks):
    n = len(tasks)
    prec = np.array([float(t[0]) for t in tasks], dtype=np.float64)
    outs = (_vp * n)(*[t[1]._h for t in tasks])
    opers = (_vp * n)(*[t[2]._h for t in tasks])
    inps = (_vp * n)(*[t[3]._h for t in tasks])
    return n, prec, outs, opers, inps


def apply_predicted(tasks, predictor=None, centers=None, collect=False):
    """apply_batch with a refinement predictor (None: plain) and per-task
    source centres (a list of [k, 3] arrays or None).  collect=True also
    returns the training samples (features [N, 12], labels [N]) of every
    output node beyond the second step.  Statistics include 'n_predicted',
    'n_cropped' and 'n_fallback'."""
    n, prec, outs, opers, inps = _batch_args(tasks)
    cs = [np.zeros((0, 3)) if centers is None or centers[i] is None else np.asarray(centers[i], dtype=np.float64).reshape(-1, 3)
          for i in range(n)]
    cflat = np.ascontiguousarray(np.concatenate(cs) if n else np.zeros((0, 3)), dtype=np.float64)
    ncen = np.array([len(c) for c in cs], dtype=np.int32)
    st = np.zeros(16 * max(n, 1))
    _check(lib().mrcpp_apply_batch_predicted(n, _p(prec), outs, opers, inps, predictor._h if predictor is not None else None,
                                             _p(cflat), _p(ncen, _ip), int(bool(collect)), _p(st)))
    stats = []
    for i in range(n):
        d = {k: float(v) for k, v in zip(STAT_KEYS[:13], st[16 * i:16 * i + 13])}
        d.update(n_predicted=st[16 * i + 13], n_cropped=st[16 * i + 14], n_fallback=st[16 * i + 15])
        stats.append(d)
    if not collect:
        return stats
    m = lib().mrcpp_samples_count()
    x = np.zeros((m, PREDICTOR_FEATURES), dtype=np.float32)
    y = np.zeros(m, dtype=np.float32)
    if lib().mrcpp_samples_fetch(_p(x, _fp), _p(y, _fp), m) != m:
        raise MRCPPError("sample fetch failed")
    return stats, x, y


def apply_batch(tasks):
    """Many applications at once: tasks is a sequence of (prec, out, oper, inp)
    or (prec, out, oper, inp, maxIter, absPrec).  Every breadth-first step of
    all trees shares one screen and one band launch; each output tree equals
    what apply() alone produces.  Returns one statistics dict per task
    ('seconds' is the whole batch's wall time)."""
    n = len(tasks)
    prec = np.array([float(t[0]) for t in tasks], dtype=np.float64)
    outs = (_vp * n)(*[t[1]._h for t in tasks])
    opers = (_vp * n)(*[t[2]._h for t in tasks])
    inps = (_vp * n)(*[t[3]._h for t in tasks])
    mi = np.array([int(t[4]) if len(t) > 4 else -1 for t in tasks], dtype=np.int32)
    ap = np.array([int(bool(t[5])) if len(t) > 5 else 0 for t in tasks], dtype=np.int32)
    st = np.zeros(16 * max(n, 1))
    _check(lib().mrcpp_apply_batch(n, _p(prec), outs, opers, inps, _p(mi, _ip), _p(ap, _ip), _p(st)))
    return [{k: float(v) for k, v in zip(STAT_KEYS, st[16 * i:16 * i + 16])} for i in range(n)]


class CompFunction:
    """CompFunction<3> with up to 4 real components (FunctionTree objects)."""

    def __init__(self, mra, ncomp=1):
        self.mra = mra
        self.CompD = [FunctionTree(mra) if i < ncomp else None for i in range(4)]

    def Ncomp(self):
        n = 0
        for i, t in enumerate(self.CompD):
            if t is not None:
                n = i + 1
        return n


def _apply_comp(prec, out, oper, inp, metric, maxIter, absPrec):
    """apply over CompFunctions (src/treebuilders/apply.cpp:119-139): out gets
    inp's component count, metric a real 4x4 [input][output] (identity)."""
    m = np.eye(4) if metric is None else np.ascontiguousarray(metric, dtype=np.float64).reshape(4, 4)
    nc = inp.Ncomp()
    out.CompD = [FunctionTree(out.mra) if i < nc else None for i in range(4)]
    for ic in range(nc):
        for oc in range(4):
            if m[ic, oc] ** 2 > 1e-15 and out.CompD[oc] is None:
                out.CompD[oc] = FunctionTree(out.mra)
    oh = (_vp * 4)(*[t._h if t is not None else None for t in out.CompD])
    ih = (_vp * 4)(*[t._h if t is not None else None for t in inp.CompD])
    _check(lib().mrcpp_apply_components(float(prec), oh, oper._h, ih, nc, _p(np.ascontiguousarray(m.ravel())), int(maxIter), int(bool(absPrec))))


def apply(prec, out, oper, inp, maxIter=-1, absPrec=False, precTrees=None, metric=None):
    """apply(prec, out, oper, inp, maxIter, absPrec), or with precision trees
    apply(prec, out, oper, inp, precTrees, maxIter, absPrec) as the reference's
    overload (src/treebuilders/apply.h:43): a list in the fifth place is
    precTrees.  With CompFunction arguments the component overload
    (apply.h:42; metric= a real 4x4 array).  Returns the run statistics
    (None for CompFunctions)."""
    if isinstance(out, CompFunction):
        return _apply_comp(prec, out, oper, inp, metric, maxIter, absPrec)
    if isinstance(maxIter, (list, tuple)):
        # positional form of the overload: (precTrees, maxIter, absPrec) in places 5-7
        trees = list(maxIter)
        maxIter = -1 if absPrec is False else int(absPrec)
        absPrec = bool(precTrees) if precTrees is not None else False
        precTrees = trees
    st = np.zeros(16)
    if precTrees is None:
        _check(lib().mrcpp_apply(float(prec), out._h, oper._h, inp._h, int(maxIter), int(bool(absPrec)), _p(st)))
    else:
        pts = (_vp * max(1, len(precTrees)))(*[t._h for t in precTrees])
        _check(lib().mrcpp_apply_prec(float(prec), out._h, oper._h, inp._h, pts, len(precTrees), int(maxIter), int(bool(absPrec)),
                                      _p(st)))
    stats = {k: float(v) for k, v in zip(STAT_KEYS, st)}
    lv = np.zeros(128, dtype=np.int32)
    n = lib().mrcpp_last_apply_levels(_p(lv, _ip), 128)
    stats["level_nodes"] = lv[:n].tolist()
    return stats
