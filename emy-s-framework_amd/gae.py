"""Generalized advantage estimation on the device (include/rsx_gae.h): the
mirror of the reference's rsc::ComputeGAEOnDevice
(RocketSimCuda/include/RocketSimCudaTraining.h:21-34) over torch tensors on
the GPU.  No CPU path: the library must be built and the tensors must live on
the HIP device."""
import ctypes as C

from .batch import _check, load_library


class GAEParams(C.Structure):
    """DeviceGAEParams (RocketSimCudaTraining.h:7-12)."""
    _fields_ = [("gamma", C.c_float), ("lam", C.c_float), ("return_std", C.c_float), ("clip_range", C.c_float)]


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def ComputeGAEOnDevice(rews, terminals, val_preds, trunc_val_preds, gamma=0.99, lam=0.95, return_std=0.0,
                       clip_range=0.0, stream=None):
    """rews, val_preds: float32 [T]; terminals: int8 [T] (0 none, 1 normal,
    2 truncated); trunc_val_preds: float32 [number of truncated steps] in
    buffer order (or None).  Returns (advantages, target_values, returns,
    rew_clip_portion as a 1-element tensor), asynchronously on `stream`
    (default: torch's current stream)."""
    import torch
    lib = load_library()
    T = int(rews.numel())
    for t in (rews, terminals, val_preds):
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError("GAE inputs must be contiguous tensors on the GPU")
    if terminals.dtype != torch.int8 or rews.dtype != torch.float32 or val_preds.dtype != torch.float32:
        raise TypeError("GAE expects float32 rewards / values and int8 terminals")
    n_tr = int(trunc_val_preds.numel()) if trunc_val_preds is not None else 0
    adv = torch.empty_like(rews)
    target = torch.empty_like(rews)
    ret = torch.empty_like(rews)
    clip = torch.zeros(1, dtype=torch.float32, device=rews.device)
    st = stream if stream is not None else torch.cuda.current_stream(rews.device).cuda_stream
    p = GAEParams(gamma, lam, return_std, clip_range)
    _check(lib.rsx_gae(_ptr(rews), _ptr(terminals), _ptr(val_preds), _ptr(trunc_val_preds), n_tr, _ptr(adv),
                       _ptr(target), _ptr(ret), _ptr(clip), T, C.byref(p), C.c_void_p(st)))
    return adv, target, ret, clip
