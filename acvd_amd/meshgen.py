"""Seeded synthetic surfaces of the clustering workloads (host C++ generator).

  icosphere(level, seed, sigma)   unit icosphere, radially displaced by N(0, sigma)
  star_icosphere(level, seed)     icosphere with a random spherical-harmonic radius
  cube(n, seed, jitter)           [-1,1]^3 with n x n quads per face, tangential jitter
  torus(nu, nv, seed)             torus R=1, r=0.35 + 32 seeded sinusoids
  mean_curvature(xyz, tri)        |H| by the cotangent Laplacian (FP64)
"""
import ctypes

import numpy as np

from ._native import meshgen, AcvdError


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def icosphere(level, seed=1, sigma=0.0):
    G = meshgen()
    nv, nf = ctypes.c_int64(), ctypes.c_int64()
    if G.acvd_gen_icosphere_sizes(level, ctypes.byref(nv), ctypes.byref(nf)):
        raise AcvdError("bad icosphere level")
    xyz = np.empty((nv.value, 3), np.float64)
    tri = np.empty((nf.value, 3), np.int32)
    G.acvd_gen_icosphere(level, seed, sigma, _p(xyz), _p(tri))
    return xyz, tri


def star_icosphere(level, seed, amp=0.05, rmin=0.2):
    G = meshgen()
    nv, nf = ctypes.c_int64(), ctypes.c_int64()
    if G.acvd_gen_icosphere_sizes(level, ctypes.byref(nv), ctypes.byref(nf)):
        raise AcvdError("bad icosphere level")
    xyz = np.empty((nv.value, 3), np.float32)
    tri = np.empty((nf.value, 3), np.int32)
    if G.acvd_gen_star_icosphere(level, seed, amp, rmin, _p(xyz), _p(tri)) < 0:
        raise AcvdError("star icosphere generation failed")
    return xyz, tri


def cube(n=400, seed=2, jitter=1e-4):
    G = meshgen()
    nv, nf = ctypes.c_int64(), ctypes.c_int64()
    if G.acvd_gen_cube_sizes(n, ctypes.byref(nv), ctypes.byref(nf)):
        raise AcvdError("bad cube resolution")
    xyz = np.empty((nv.value, 3), np.float32)
    tri = np.empty((nf.value, 3), np.int32)
    if G.acvd_gen_cube(n, seed, jitter, _p(xyz), _p(tri)):
        raise AcvdError("cube generation failed")
    return xyz, tri


def torus(nu=4096, nv=2048, seed=3, R=1.0, r0=0.35, nsin=32, amp=0.01, fmax=32):
    G = meshgen()
    V, F = ctypes.c_int64(), ctypes.c_int64()
    if G.acvd_gen_torus_sizes(nu, nv, ctypes.byref(V), ctypes.byref(F)):
        raise AcvdError("bad torus resolution")
    xyz = np.empty((V.value, 3), np.float32)
    tri = np.empty((F.value, 3), np.int32)
    G.acvd_gen_torus(nu, nv, R, r0, nsin, amp, fmax, seed, _p(xyz), _p(tri))
    return xyz, tri


def mean_curvature(xyz, tri):
    G = meshgen()
    xyz = np.ascontiguousarray(xyz)
    tri = np.ascontiguousarray(tri, np.int32)
    out = np.empty(len(xyz), np.float64)
    if xyz.dtype == np.float32:
        G.acvd_mean_curvature(len(xyz), _p(xyz), None, len(tri), _p(tri), _p(out))
    else:
        x64 = np.ascontiguousarray(xyz, np.float64)
        G.acvd_mean_curvature(len(xyz), None, _p(x64), len(tri), _p(tri), _p(out))
    return out


def seed_labels(nv, K, seed):
    """K distinct seed items labelled 0..K-1 (partial Fisher-Yates), others K."""
    out = np.empty(nv, np.int32)
    if meshgen().acvd_gen_seed_labels(nv, K, seed, _p(out)):
        raise AcvdError("bad seed label request")
    return out


def batch_levels(count=384, lmin=4, lmax=8, seed=4):
    out = np.empty(count, np.int32)
    meshgen().acvd_gen_batch_levels(count, lmin, lmax, seed, _p(out))
    return out
