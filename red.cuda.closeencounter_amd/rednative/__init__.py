"""rednative -- Python side of the MI355X-native red.cuda integration path.

* ``files``: the reference's body / parameter / result / event file formats;
* ``mainloop``: the reference driver loop over any engine;
* ``native``: ctypes binding of librednative.so (HIP kernels behind a C ABI);
* ``generate``: seeded synthetic disks for the benchmark configurations.
"""
from . import files, mainloop  # noqa: F401

__all__ = ["files", "mainloop", "native", "generate"]
