"""ctypes binding of the CPU oracle (oracle/mrcpp_oracle.cpp).

Test infrastructure: the oracle restates MRCPP's FP64 apply path on the CPU so
that the MI355X implementation can be checked against it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")
FILTER_DIR = os.path.join(os.path.dirname(HERE), "mrcpp_amd", "share", "mwfilters")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle with its Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


class Oracle:
    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError("oracle library missing: run make -C oracle")
            L = ctypes.CDLL(LIB_PATH)
            L.orc_last_error.restype = ctypes.c_char_p
            L.orc_set_filter_dir.argtypes = [ctypes.c_char_p]
            L.orc_mra_create.restype = ctypes.c_void_p
            L.orc_mra_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
            L.orc_mra_destroy.argtypes = [ctypes.c_void_p]
            L.orc_mra_info.argtypes = [ctypes.c_void_p, _ip, _dp, _ip]
            L.orc_tree_create.restype = ctypes.c_void_p
            L.orc_tree_create.argtypes = [ctypes.c_void_p]
            L.orc_tree_destroy.argtypes = [ctypes.c_void_p]
            L.orc_tree_project_gaussians.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp]
            L.orc_tree_build_grid_gaussians.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
            L.orc_tree_num_nodes.restype = ctypes.c_longlong
            L.orc_tree_num_nodes.argtypes = [ctypes.c_void_p]
            L.orc_tree_square_norm.restype = ctypes.c_double
            L.orc_tree_square_norm.argtypes = [ctypes.c_void_p]
            L.orc_tree_export.argtypes = [ctypes.c_void_p, _ip, _dp, _dp]
            L.orc_tree_evalf.restype = ctypes.c_double
            L.orc_arith.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_void_p),
                                    ctypes.c_int, ctypes.c_int]
            L.orc_tree_save_txt.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
            L.orc_tree_load_txt.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
            L.orc_tree_evalf.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double]
            L.orc_tree_dot.restype = ctypes.c_double
            L.orc_tree_dot.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            L.orc_tree_integrate.restype = ctypes.c_double
            L.orc_tree_integrate.argtypes = [ctypes.c_void_p]
            L.orc_operator_create.restype = ctypes.c_void_p
            L.orc_operator_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double]
            L.orc_operator_destroy.argtypes = [ctypes.c_void_p]
            L.orc_operator_size.argtypes = [ctypes.c_void_p]
            L.orc_operator_build_ms.restype = ctypes.c_double
            L.orc_operator_build_ms.argtypes = [ctypes.c_void_p]
            L.orc_operator_kernel.argtypes = [ctypes.c_void_p, _dp, _dp]
            L.orc_kernel_terms.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, _dp, _dp, ctypes.c_int]
            L.orc_operator_bandwidths.argtypes = [ctypes.c_void_p, ctypes.c_double, _ip, ctypes.c_int]
            L.orc_operator_node.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
            L.orc_apply.argtypes = [ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int, ctypes.c_int, _dp, _ip, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]
            L.orc_set_threads.argtypes = [ctypes.c_int]
            L.orc_prec_factor.restype = ctypes.c_double
            L.orc_prec_factor.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int, _ip]
            L.orc_set_filter_dir(FILTER_DIR.encode())
            cls._lib = L
        return cls._lib

    @classmethod
    def error(cls):
        return cls.lib().orc_last_error().decode()

    @classmethod
    def set_threads(cls, n):
        cls.lib().orc_set_threads(int(n))

    @classmethod
    def max_threads(cls):
        return cls.lib().orc_max_threads()

    @classmethod
    def kernel_terms(cls, kind, mu, eps, r_min, r_max):
        a = np.zeros(2000)
        c = np.zeros(2000)
        n = cls.lib().orc_kernel_terms(kind, mu, eps, r_min, r_max, _ptr(a), _ptr(c), 2000)
        if n < 0:
            raise RuntimeError(cls.error())
        return a[:n].copy(), c[:n].copy()


class OracleMRA:
    def __init__(self, box, order, depth=30):
        L = Oracle.lib()
        self.h = L.orc_mra_create(int(box[0]), int(box[1]), int(order), int(depth))
        if not self.h:
            raise RuntimeError(Oracle.error())
        self.order = order
        n0 = ctypes.c_int()
        sf = ctypes.c_double()
        ms = ctypes.c_int()
        L.orc_mra_info(self.h, ctypes.byref(n0), ctypes.byref(sf), ctypes.byref(ms))
        self.root_scale, self.scaling_factor, self.max_scale = n0.value, sf.value, ms.value

    def __del__(self):
        if getattr(self, "h", None):
            Oracle.lib().orc_mra_destroy(self.h)
            self.h = None


class OracleTree:
    def __init__(self, mra):
        self.mra = mra
        self.h = Oracle.lib().orc_tree_create(mra.h)

    def __del__(self):
        if getattr(self, "h", None):
            Oracle.lib().orc_tree_destroy(self.h)
            self.h = None

    def project_gaussians(self, prec, coefs, alphas, centers):
        coefs = np.ascontiguousarray(coefs, dtype=np.float64)
        alphas = np.ascontiguousarray(alphas, dtype=np.float64)
        centers = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        rc = Oracle.lib().orc_tree_project_gaussians(self.h, prec, len(coefs), _ptr(coefs), _ptr(alphas), _ptr(centers))
        if rc != 0:
            raise RuntimeError(Oracle.error())

    def build_grid_gaussians(self, coefs, alphas, centers):
        coefs = np.ascontiguousarray(coefs, dtype=np.float64)
        alphas = np.ascontiguousarray(alphas, dtype=np.float64)
        centers = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        rc = Oracle.lib().orc_tree_build_grid_gaussians(self.h, len(coefs), _ptr(coefs), _ptr(alphas), _ptr(centers))
        if rc != 0:
            raise RuntimeError(Oracle.error())

    def save_txt(self, path):
        """FunctionTree::saveTreeTXT restated (MADNESS-convention text)."""
        if Oracle.lib().orc_tree_save_txt(self.h, os.fsencode(path)) != 0:
            raise RuntimeError(Oracle.error())

    def load_txt(self, path):
        """FunctionTree::loadTreeTXT restated."""
        if Oracle.lib().orc_tree_load_txt(self.h, os.fsencode(path)) != 0:
            raise RuntimeError(Oracle.error())

    def num_nodes(self):
        return Oracle.lib().orc_tree_num_nodes(self.h)

    def square_norm(self):
        return Oracle.lib().orc_tree_square_norm(self.h)

    def export(self, with_coefs=True):
        """Return (idx[N,5] = scale,lx,ly,lz,is_end; coefs[N,8*(k+1)^3]; norms[N,9])."""
        n = self.num_nodes()
        kp1_d = (self.mra.order + 1) ** 3
        idx = np.zeros((n, 5), dtype=np.int32)
        coefs = np.zeros((n, 8 * kp1_d)) if with_coefs else None
        norms = np.zeros((n, 9))
        Oracle.lib().orc_tree_export(self.h, _ptr(idx, _ip), _ptr(coefs) if with_coefs else None, _ptr(norms))
        return idx, coefs, norms

    def evalf(self, r):
        return Oracle.lib().orc_tree_evalf(self.h, float(r[0]), float(r[1]), float(r[2]))

    def dot(self, other):
        return Oracle.lib().orc_tree_dot(self.h, other.h)

    def integrate(self):
        return Oracle.lib().orc_tree_integrate(self.h)


class OracleOperator:
    POISSON, HELMHOLTZ = 0, 1

    def __init__(self, mra, prec, mu=None):
        kind = self.POISSON if mu is None else self.HELMHOLTZ
        self.mra = mra
        self.h = Oracle.lib().orc_operator_create(mra.h, kind, 0.0 if mu is None else float(mu), float(prec))
        if not self.h:
            raise RuntimeError(Oracle.error())

    def __del__(self):
        if getattr(self, "h", None):
            Oracle.lib().orc_operator_destroy(self.h)
            self.h = None

    def size(self):
        return Oracle.lib().orc_operator_size(self.h)

    def build_ms(self):
        return Oracle.lib().orc_operator_build_ms(self.h)

    def kernel(self):
        m = self.size()
        a = np.zeros(m)
        c = np.zeros(m)
        Oracle.lib().orc_operator_kernel(self.h, _ptr(a), _ptr(c))
        return a, c

    def bandwidths(self, prec, max_depth=32):
        out = np.zeros((self.size(), max_depth, 5), dtype=np.int32)
        Oracle.lib().orc_operator_bandwidths(self.h, float(prec), _ptr(out, _ip), max_depth)
        return out

    def node(self, term, depth, l, comp):
        kp1 = self.mra.order + 1
        mat = np.zeros(kp1 * kp1)
        nrm = ctypes.c_double()
        rc = Oracle.lib().orc_operator_node(self.h, term, depth, l, comp, _ptr(mat), ctypes.byref(nrm))
        if rc != 0:
            return None, None
        return mat.reshape(kp1, kp1, order="F"), nrm.value


def apply(prec, out, oper, inp, max_iter=-1, abs_prec=False, prec_trees=None):
    """Run the oracle apply (with precision trees when prec_trees is a list,
    src/treebuilders/apply.cpp:214-251); returns a stats dict."""
    stats = np.zeros(10)
    levels = np.zeros(64, dtype=np.int32)
    pts = (ctypes.c_void_p * max(1, len(prec_trees or ())))(*[t.h for t in (prec_trees or ())])
    rc = Oracle.lib().orc_apply(float(prec), out.h, oper.h, inp.h, int(max_iter), int(abs_prec),
                                _ptr(stats), _ptr(levels, _ip), 64, pts, -1 if prec_trees is None else len(prec_trees))
    if rc != 0:
        raise RuntimeError(Oracle.error())
    nlev = int(stats[4])
    return {"n_applied": int(stats[0]), "n_gen_applied": int(stats[1]), "n_g": int(stats[2]),
            "seconds": float(stats[3]), "level_nodes": levels[:nlev].tolist(),
            "n_band_pairs": int(stats[5]), "n_pairs": int(stats[6]), "n_groups": int(stats[7]),
            "n_xsweeps": int(stats[8]), "n_ysweeps": int(stats[9])}


def prec_factor(trees, n, l):
    """precFunc of apply with precision trees at node (n, l)
    (src/treebuilders/apply.cpp:222-229)."""
    pts = (ctypes.c_void_p * max(1, len(trees)))(*[t.h for t in trees])
    la = np.array(l, dtype=np.int32)
    v = Oracle.lib().orc_prec_factor(pts, len(trees), int(n), _ptr(la, _ip))
    if v < 0:
        raise RuntimeError(Oracle.error())
    return v


def _arith(mode, prec, out, terms, max_iter, abs_prec):
    c = np.ascontiguousarray([float(a) for a, _ in terms], dtype=np.float64)
    hs = (ctypes.c_void_p * len(terms))(*[t.h for _, t in terms])
    if Oracle.lib().orc_arith(mode, prec, out.h, len(terms), _ptr(c), hs, max_iter, int(abs_prec)) != 0:
        raise RuntimeError(Oracle.error())


def add(prec, out, terms, max_iter=-1, abs_prec=False):
    """add(prec, out, [(c_i, tree_i), ...]) restated (src/treebuilders/add.cpp:43-80)."""
    _arith(0, prec, out, terms, max_iter, abs_prec)


def multiply(prec, out, terms, max_iter=-1, abs_prec=False, use_max_norms=False):
    """multiply(prec, out, [(c_i, tree_i), ...]) restated (src/treebuilders/multiply.cpp:52-119);
    use_max_norms refines with MultiplicationAdaptor (two trees)."""
    _arith(2 if use_max_norms else 1, prec, out, terms, max_iter, abs_prec)
