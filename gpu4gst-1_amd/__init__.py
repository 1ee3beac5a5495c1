"""MI355X-native group Steiner tree (GST) engine.

The directory name is not a Python identifier, so the package is loaded by path under
the module name ``gpu4gst_1_amd`` (see ``load_package`` in the repository's
``__graft_entry__.py``).

Submodules:
  datasets  seeded synthetic workloads and the reference's on-disk formats
  engine    ctypes mirror of the C-ABI in include/gst.h (libgst.so, HIP for gfx950)
  build     compiles libgst.so / libgstdata.so in-tree
"""
