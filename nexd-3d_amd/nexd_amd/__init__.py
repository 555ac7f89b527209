"""NEXD-3D time-stepping engine for MI355X: host-side set-up (reference
element, mesh, materials, sources/receivers), the reference's file formats,
and the ctypes binding of the C-ABI library ``libnexd_gpu.so``."""
