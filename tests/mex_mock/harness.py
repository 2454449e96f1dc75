"""Python side of the MEX stand-in: converts numpy arrays, lists (cells)
and scipy sparse matrices to mxArrays, runs a shim's mexFunction through
mexmock_call, and converts the outputs back (MATLAB column-major order).
Handle structs returned by the shims come back as ``MxRef`` and can be
passed to later calls.  Test infrastructure only."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, '_build')

mx_p = C.c_void_p
_CLS_DOUBLE, _CLS_UINT64 = 4, 5


class MexError(RuntimeError):
    pass


class MxRef:
    """An mxArray owned by the harness (handle structs)."""

    def __init__(self, ptr):
        self.ptr = ptr


class Mex:
    def __init__(self):
        import build_mock
        if not os.path.exists(os.path.join(OUT, 'libmexmock.so')) or any(
                not os.path.exists(os.path.join(OUT, n + '.so')) for n in build_mock.shims()):
            build_mock.build()
        self.m = C.CDLL(os.path.join(OUT, 'libmexmock.so'), mode=C.RTLD_GLOBAL)
        m = self.m
        sz = C.c_size_t
        for name, res, args in [
            ('mxCreateNumericArray', mx_p, [sz, C.POINTER(sz), C.c_int, C.c_int]),
            ('mxCreateCellMatrix', mx_p, [sz, sz]),
            ('mxCreateSparse', mx_p, [sz, sz, sz, C.c_int]),
            ('mxCreateString', mx_p, [C.c_char_p]),
            ('mxSetCell', None, [mx_p, sz, mx_p]),
            ('mxGetCell', mx_p, [mx_p, sz]),
            ('mxGetData', C.c_void_p, [mx_p]),
            ('mxGetIr', C.c_void_p, [mx_p]),
            ('mxGetJc', C.c_void_p, [mx_p]),
            ('mxGetNumberOfDimensions', sz, [mx_p]),
            ('mxGetDimensions', C.POINTER(sz), [mx_p]),
            ('mxGetNumberOfElements', sz, [mx_p]),
            ('mxIsCell', C.c_int, [mx_p]),
            ('mxIsStruct', C.c_int, [mx_p]),
            ('mxIsComplex', C.c_int, [mx_p]),
            ('mxIsUint64', C.c_int, [mx_p]),
            ('mxDuplicateArray', mx_p, [mx_p]),
            ('mxDestroyArray', None, [mx_p]),
            ('mexmock_call', C.c_int, [C.c_void_p, C.c_int, C.POINTER(mx_p), C.c_int, C.POINTER(mx_p), C.c_char_p,
                                       C.c_int]),
            ('mexmock_run_at_exit', None, []),
            ('mxGetString', C.c_int, [mx_p, C.c_char_p, sz]),
        ]:
            f = getattr(m, name)
            f.restype = res
            f.argtypes = args
        self._shims = {}
        # MATLAB functions the shims call back: wfilters(name, 'r') -> the
        # reconstruction low-pass filter (fhs.sara's Daubechies construction)
        proto = C.CFUNCTYPE(C.c_int, C.c_int, C.POINTER(mx_p), C.c_int, C.POINTER(mx_p), C.c_char_p)

        def matlab(nlhs, plhs, nrhs, prhs, name):
            try:
                if name.decode() != 'wfilters' or nrhs < 1:
                    return 1
                buf = C.create_string_buffer(64)
                m.mxGetString(prhs[0], buf, 64)
                wname = buf.value.decode()
                if not wname.startswith('db'):
                    return 1
                from fhs.sara import daubechies_scaling_filter
                f = daubechies_scaling_filter(int(wname[2:]))
                plhs[0] = self.to_mx(f[None, :])
                return 0
            except Exception:
                return 1
        self._matlab = proto(matlab)      # kept alive with the harness
        m.mexmock_set_matlab.argtypes = [proto]
        m.mexmock_set_matlab.restype = None
        m.mexmock_set_matlab(self._matlab)

    # ---- conversions ----------------------------------------------------
    def to_mx(self, obj):
        m = self.m
        if isinstance(obj, MxRef):
            return m.mxDuplicateArray(obj.ptr)
        if isinstance(obj, (list, tuple)):
            c = m.mxCreateCellMatrix(len(obj), 1)
            for i, o in enumerate(obj):
                m.mxSetCell(c, i, self.to_mx(o))
            return c
        if isinstance(obj, str):
            return m.mxCreateString(obj.encode())
        if hasattr(obj, 'tocsc'):
            A = obj.tocsc()
            A.sort_indices()
            cplx = np.iscomplexobj(A.data)
            s = m.mxCreateSparse(A.shape[0], A.shape[1], max(A.nnz, 1), int(cplx))
            # the converted arrays are bound to names: a temporary's buffer
            # would be freed before memmove reads it
            jc = np.ascontiguousarray(A.indptr, dtype=np.uint64)
            C.memmove(m.mxGetJc(s), jc.ctypes.data, jc.nbytes)
            if A.nnz:
                ir = np.ascontiguousarray(A.indices, dtype=np.uint64)
                C.memmove(m.mxGetIr(s), ir.ctypes.data, ir.nbytes)
                d = np.ascontiguousarray(A.data, dtype=np.complex128 if cplx else np.float64)
                C.memmove(m.mxGetData(s), d.ctypes.data, d.nbytes)
            return s
        a = np.asarray(obj)
        cplx = np.iscomplexobj(a)
        dims = a.shape if a.ndim >= 2 else ((a.size, 1) if a.ndim == 1 else (1, 1))
        d = (C.c_size_t * len(dims))(*dims)
        mxa = m.mxCreateNumericArray(len(dims), d, _CLS_DOUBLE, int(cplx))
        buf = np.ascontiguousarray(np.asarray(a, dtype=np.complex128 if cplx else np.float64).ravel(order='F'))
        if buf.size:
            C.memmove(m.mxGetData(mxa), buf.ctypes.data, buf.nbytes)
        return mxa

    def from_mx(self, p):
        m = self.m
        if m.mxIsStruct(p):
            return MxRef(p)
        if m.mxIsCell(p):
            n = m.mxGetNumberOfElements(p)
            out = [self.from_mx(m.mxDuplicateArray(m.mxGetCell(p, i))) for i in range(n)]
            m.mxDestroyArray(p)
            return out
        nd = m.mxGetNumberOfDimensions(p)
        dp = m.mxGetDimensions(p)
        dims = tuple(dp[i] for i in range(nd))
        n = int(np.prod(dims))
        dt = np.complex128 if m.mxIsComplex(p) else (np.uint64 if m.mxIsUint64(p) else np.float64)
        arr = np.empty(n, dtype=dt)
        if n:
            C.memmove(arr.ctypes.data, m.mxGetData(p), arr.nbytes)
        m.mxDestroyArray(p)
        arr = arr.reshape(dims, order='F')
        return arr.item() if arr.size == 1 and dims == (1, 1) else arr

    # ---- calls ----------------------------------------------------------
    def call(self, name, nlhs, *args):
        """[out1, ..., out_nlhs] = name(args...); raises MexError."""
        lib = self._shims.get(name)
        if lib is None:
            lib = self._shims[name] = C.CDLL(os.path.join(OUT, name + '.so'))
        fn = C.cast(lib.mexFunction, C.c_void_p)
        prhs = (mx_p * max(1, len(args)))(*[self.to_mx(a) for a in args])
        plhs = (mx_p * max(1, nlhs))()
        err = C.create_string_buffer(1024)
        rc = self.m.mexmock_call(fn, nlhs, plhs, len(args), prhs, err, len(err))
        for i in range(len(args)):
            self.m.mxDestroyArray(prhs[i])
        if rc:
            for i in range(nlhs):
                if plhs[i]:
                    self.m.mxDestroyArray(plhs[i])
            raise MexError(err.value.decode())
        outs = [self.from_mx(plhs[i]) for i in range(nlhs)]
        return outs[0] if nlhs == 1 else outs

    def close(self):
        self.m.mexmock_run_at_exit()
