"""torch.autograd binding of the filter-convolve-window-integrate stage.

`filter_convolve_window(tf, fw, bw)` evaluates the base kernels of every kernel at every
point from the filter responses and the strain traces, as the receiver loop of
slave_work does (src/slave.f90:515-534, 586-640; kernel.f90:410-570, 946-1020), through
mck_ad_forward of include/mck.h; its backward is the hand-written adjoint kernel behind
mck_ad_backward, which recomputes the FFTs from the traces.  This binding exists for
the gradient check of the stage; the production path is mck_slave_work.

Shapes (nd = 1 for trace-only kernels, 6 for full strain; nomega = N/2 + 1):
  tf   complex64 (nfilter, nomega)   filter responses G; TF_fwd TF_bwd = G^2 x set-up factors
  fw   float32   (npts, nd, nt)      earth-frame fwd strain traces
  bw   float32   (npts, nrec, nd, nt) earth-frame bwd strain traces
  ->   float64   (npts, nkernel, 6)  base kernels lambda, mu, rho, a, b, c
Gradients of tf follow torch's convention for complex inputs (d/dRe + i d/dIm of a real
loss).  All tensors live on the GPU; the kernels run on torch's current stream.
"""
import ctypes as C

import numpy as np
import torch

from . import _check, load_library


def dims():
    """(nd, nt, nomega, nfilter, nrec, nkernel) of the initialised set-up."""
    v = [C.c_int32() for _ in range(6)]
    _check(load_library().mck_ad_dims(*[C.byref(x) for x in v]))
    return tuple(int(x.value) for x in v)


def transfer_functions(device="cuda"):
    """The set-up's own filter responses G as a complex64 tensor (nfilter, nomega)."""
    nd, nt, nw, nf, nr, nk = dims()
    host = np.zeros((nf, nw, 2), np.float32)
    _check(load_library().mck_ad_transfer(host.ctypes.data_as(C.POINTER(C.c_float))))
    return torch.view_as_complex(torch.from_numpy(host)).to(device)


def _addr(t):
    return C.c_void_p(t.data_ptr())


def _stream(t):
    return C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _check_inputs(tf, fw, bw):
    nd, nt, nw, nf, nr, nk = dims()
    if tf.dtype != torch.complex64 or tuple(tf.shape) != (nf, nw):
        raise ValueError("tf must be complex64 of shape (%d, %d)" % (nf, nw))
    if fw.dtype != torch.float32 or fw.dim() != 3 or tuple(fw.shape[1:]) != (nd, nt):
        raise ValueError("fw must be float32 of shape (npts, %d, %d)" % (nd, nt))
    if bw.dtype != torch.float32 or tuple(bw.shape) != (fw.shape[0], nr, nd, nt):
        raise ValueError("bw must be float32 of shape (%d, %d, %d, %d)" % (fw.shape[0], nr, nd, nt))
    for t in (tf, fw, bw):
        if not t.is_cuda:
            raise ValueError("the stage runs on the GPU only: inputs must be device tensors")
    return nk


class FilterConvolveWindow(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tf, fw, bw):
        nk = _check_inputs(tf, fw, bw)
        tf, fw, bw = tf.contiguous(), fw.contiguous(), bw.contiguous()
        base = torch.empty((fw.shape[0], nk, 6), dtype=torch.float64, device=fw.device)
        if fw.shape[0]:
            _check(load_library().mck_ad_forward(fw.shape[0], _addr(tf), _addr(fw), _addr(bw),
                                                 _addr(base), _stream(fw)))
        ctx.save_for_backward(tf, fw, bw)
        return base

    @staticmethod
    def backward(ctx, gbase):
        tf, fw, bw = ctx.saved_tensors
        gbase = gbase.to(torch.float64).contiguous()
        gtf = torch.empty_like(tf)
        gfw = torch.empty_like(fw)
        gbw = torch.empty_like(bw)
        _check(load_library().mck_ad_backward(fw.shape[0], _addr(tf), _addr(fw), _addr(bw),
                                              _addr(gbase), _addr(gtf), _addr(gfw), _addr(gbw),
                                              _stream(fw)))
        return gtf, gfw, gbw


def filter_convolve_window(tf, fw, bw):
    return FilterConvolveWindow.apply(tf, fw, bw)
