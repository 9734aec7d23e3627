"""MI355X-native HashDAG engine: device-resident hash-consed voxel DAG, VBR colors, GC and
ESVO ray traversal, driven through the C-ABI in include/hd_capi.h (see vkhashdag_amd.capi)."""
