"""MI355X MDFT functional engine: host side.

`mdft.gpu.Engine` drives the HIP library libmdft_gpu.so through its C-ABI
(include/mdft_gpu.h); `mdft.ck` reads, writes and generates c(k) tables;
`mdft.constants` holds the solvent model and physical constants.
"""
from . import ck, constants  # noqa: F401
