"""Host-side helpers of the MI355X SENGA+ right-hand-side library.

Readers and writers for the reference's input formats (cont.dat, chem.dat,
diff.dat, restart dumps), seeded initial-condition builders, and the ctypes
binding of the C ABI in include/senga_gpu.h.
"""
