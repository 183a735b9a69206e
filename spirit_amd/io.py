"""Spin configurations and chains in OVF 2.0 files; mirrors the reference's ``spirit.io``
(core/python/spirit/io.py) for the OVF entry points of ``IO.h``."""
import ctypes

from spirit_amd import spiritlib

_spirit = spiritlib.load_spirit_library()

FILEFORMAT_OVF_BIN = 0
FILEFORMAT_OVF_BIN4 = 1
FILEFORMAT_OVF_BIN8 = 2
FILEFORMAT_OVF_TEXT = 3
FILEFORMAT_OVF_CSV = 4

_vp, _i, _s = ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p


def _fn(name, args, res=None):
    f = getattr(_spirit, name)
    f.argtypes = args
    f.restype = res
    return f


_N_Images_In_File = _fn("IO_N_Images_In_File", [_vp, _s, _i, _i], _i)
_Image_Read = _fn("IO_Image_Read", [_vp, _s, _i, _i, _i])
_Image_Write = _fn("IO_Image_Write", [_vp, _s, _i, _s, _i, _i])
_Image_Append = _fn("IO_Image_Append", [_vp, _s, _i, _s, _i, _i])
_Chain_Read = _fn("IO_Chain_Read", [_vp, _s, _i, _i, _i, _i])
_Chain_Write = _fn("IO_Chain_Write", [_vp, _s, _i, _s, _i])
_Chain_Append = _fn("IO_Chain_Append", [_vp, _s, _i, _s, _i])
_Positions_Write = _fn("IO_Positions_Write", [_vp, _s, _i, _s, _i, _i])


def _b(s):
    return s.encode("utf-8")


def n_images_in_file(p_state, filename, idx_image_inchain=-1, idx_chain=-1):
    """Number of images (OVF segments) in a file, -1 if it is not OVF."""
    return int(_N_Images_In_File(ctypes.c_void_p(p_state), _b(filename), idx_image_inchain, idx_chain))


def image_read(p_state, filename, idx_image_infile=0, idx_image_inchain=-1, idx_chain=-1):
    """Read image idx_image_infile of the file into an image of the chain."""
    _Image_Read(ctypes.c_void_p(p_state), _b(filename), idx_image_infile, idx_image_inchain, idx_chain)


def image_write(p_state, filename, fileformat=FILEFORMAT_OVF_TEXT, comment="", idx_image=-1, idx_chain=-1):
    """Write an image of the chain to a (new) file."""
    _Image_Write(ctypes.c_void_p(p_state), _b(filename), fileformat, _b(comment), idx_image, idx_chain)


def image_append(p_state, filename, fileformat=FILEFORMAT_OVF_TEXT, comment="", idx_image=-1, idx_chain=-1):
    """Append an image of the chain to a file as one more segment."""
    _Image_Append(ctypes.c_void_p(p_state), _b(filename), fileformat, _b(comment), idx_image, idx_chain)


def chain_read(p_state, filename, starting_image=0, ending_image=-1, insert_idx=-1, idx_chain=-1):
    """Read images starting_image .. ending_image of the file into the chain from insert_idx on."""
    _Chain_Read(ctypes.c_void_p(p_state), _b(filename), starting_image, ending_image, insert_idx, idx_chain)


def chain_write(p_state, filename, fileformat=FILEFORMAT_OVF_TEXT, comment="", idx_chain=-1):
    """Write every image of the chain to a (new) file."""
    _Chain_Write(ctypes.c_void_p(p_state), _b(filename), fileformat, _b(comment), idx_chain)


def chain_append(p_state, filename, fileformat=FILEFORMAT_OVF_TEXT, comment="", idx_chain=-1):
    """Append every image of the chain to a file."""
    _Chain_Append(ctypes.c_void_p(p_state), _b(filename), fileformat, _b(comment), idx_chain)


def positions_write(p_state, filename, fileformat=FILEFORMAT_OVF_BIN, comment="-", idx_image=-1, idx_chain=-1):
    """Write the spin positions as a vector field."""
    _Positions_Write(ctypes.c_void_p(p_state), _b(filename), fileformat, _b(comment), idx_image, idx_chain)
