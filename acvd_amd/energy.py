"""Differentiable fixed-assignment clustering energy.

For frozen labels the clustering energy of each mesh is a smooth function of
the vertex positions x and weights w (isotropic: sum w |x - c|^2; QEM:
sum w |x - q|^2 with q the rank-truncated quadric point of the cluster,
the cluster quadric recomputed from the face planes).  The forward and the
hand-written adjoint run in libacvd_amd.so (k_energy.hip) on device tensors;
`ClusterEnergyFunction` exposes them to torch.autograd.
"""
import ctypes

import numpy as np
import torch

from . import _native as N
from ._native import AcvdError, check


class ClusterEnergy:
    """Energy context for one or more meshes (packed in the order added)."""

    def __init__(self, device=0):
        self.device = int(device)
        self._h = ctypes.c_void_p()
        check(N.lib().acvd_energy_create(ctypes.byref(self._h), self.device))
        self._nv = []
        self._gen = 0

    def close(self):
        if self._h:
            N.lib().acvd_energy_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_mesh(self, n_vertices, triangles, n_clusters, metric, quadrics_level=3):
        tri = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
        mid = ctypes.c_int32()
        check(N.lib().acvd_energy_add_mesh(self._h, int(n_vertices), len(tri),
                                           tri.ctypes.data_as(ctypes.c_void_p), int(n_clusters), int(metric),
                                           int(quadrics_level), ctypes.byref(mid)))
        self._nv.append(int(n_vertices))
        return mid.value

    @property
    def n_items(self):
        return sum(self._nv)

    @property
    def n_meshes(self):
        return len(self._nv)

    def _check(self, xyz, w, labels):
        n = self.n_items
        dev = torch.device("cuda", self.device)
        if xyz.device != dev or w.device != dev or labels.device != dev:
            raise AcvdError(f"inputs must be on {dev}")
        if xyz.dtype != torch.float32 or w.dtype != torch.float32 or labels.dtype != torch.int32:
            raise AcvdError("xyz and w must be float32, labels int32")
        if tuple(xyz.shape) != (n, 3) or tuple(w.shape) != (n,) or tuple(labels.shape) != (n,):
            raise AcvdError(f"expected xyz ({n}, 3), w ({n},), labels ({n},)")
        if not (xyz.is_contiguous() and w.is_contiguous() and labels.is_contiguous()):
            raise AcvdError("inputs must be contiguous")

    def forward_raw(self, xyz, w, labels):
        """Per-mesh energies (float64 device tensor); no autograd."""
        self._check(xyz, w, labels)
        e = torch.empty(self.n_meshes, dtype=torch.float64, device=xyz.device)
        stream = torch.cuda.current_stream(xyz.device).cuda_stream
        check(N.lib().acvd_energy_forward(self._h, ctypes.c_void_p(xyz.data_ptr()), ctypes.c_void_p(w.data_ptr()),
                                          ctypes.c_void_p(labels.data_ptr()), ctypes.c_void_p(e.data_ptr()),
                                          ctypes.c_void_p(stream)))
        self._gen += 1
        return e

    def backward_raw(self, xyz, w, labels, grad_energy):
        """dE/dxyz (float32, (N, 3)) and dE/dw (float32, (N,)) for the upstream
        per-mesh gradient; the last forward must have run on these inputs."""
        self._check(xyz, w, labels)
        g = grad_energy.to(device=xyz.device, dtype=torch.float64).contiguous()
        dx = torch.empty_like(xyz)
        dw = torch.empty_like(w)
        stream = torch.cuda.current_stream(xyz.device).cuda_stream
        check(N.lib().acvd_energy_backward(self._h, ctypes.c_void_p(xyz.data_ptr()), ctypes.c_void_p(w.data_ptr()),
                                           ctypes.c_void_p(labels.data_ptr()), ctypes.c_void_p(g.data_ptr()),
                                           ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(dw.data_ptr()),
                                           ctypes.c_void_p(stream)))
        return dx, dw

    def __call__(self, xyz, w, labels):
        return ClusterEnergyFunction.apply(xyz, w, labels, self)


class ClusterEnergyFunction(torch.autograd.Function):
    """E = ClusterEnergyFunction.apply(xyz, w, labels, energy_context):
    per-mesh energies (float64); gradients flow to xyz and w."""

    @staticmethod
    def forward(ctx, xyz, w, labels, energy):
        e = energy.forward_raw(xyz, w, labels)
        ctx.save_for_backward(xyz, w, labels)
        ctx.energy = energy
        ctx.gen = energy._gen
        return e

    @staticmethod
    def backward(ctx, grad_e):
        xyz, w, labels = ctx.saved_tensors
        energy = ctx.energy
        if energy._gen != ctx.gen:
            # another forward ran on this context since: restore its cluster state
            energy.forward_raw(xyz, w, labels)
            ctx.gen = energy._gen
        dx, dw = energy.backward_raw(xyz, w, labels, grad_e)
        return dx, dw, None, None
