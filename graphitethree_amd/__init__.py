"""MI355X-native OpenNL solve path: CRS SpMV, Jacobi-PCG and NLBlas as HIP
kernels for gfx950 behind OpenNL's C API (libnlhip.so), with a Python mirror
of Graphite's OGF::NL scripting classes (graphitethree_amd.nl)."""
