"""cubff on MI355X: the per-epoch hot path of the cubff primordial-soup
simulator as gfx950 HIP kernels behind a C-ABI (include/cubff_hip.h).

cubff_amd.capi binds that C-ABI with ctypes; the native library lives in
cubff_amd/lib/ and is built by __graft_entry__.build().
"""
