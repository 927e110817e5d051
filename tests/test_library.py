"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: no GPU is needed)."""
import ctypes
import glob
import os
import re

from mt_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(mte_[a-z_0-9]+)\s*\(", text))
    return names


def test_header_symbols_exported():
    lib = ctypes.CDLL(engine.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_header():
    assert declared_symbols() <= set(engine.SIGNATURES)


def test_struct_layouts_match_header():
    # mte_params: 36 4-byte fields in declaration order; mte_coord is 7 floats
    from mt_amd.params import Params
    assert ctypes.sizeof(Params) == 4 * len(Params._fields_)
    assert ctypes.sizeof(engine.Options) == 20
    assert ctypes.sizeof(engine.Topology) == 8 + 14 * 8
    assert ctypes.sizeof(engine.Tuning) == 8 * 4


def test_binding_rejects_mis_sized_batch_arrays():
    """The ctypes binding checks every per-bead / per-trajectory array against
    the batch shape before it crosses the ABI (the C side takes plain
    pointers), and passes an empty box as NULL (open boundaries)."""
    import numpy as np
    import pytest
    from mt_amd import params as P
    from mt_amd.engine import Engine
    from tests.helpers import c1_batch
    b = c1_batch(ntraj=2)
    b.gtp = b.gtp[:-1]
    with pytest.raises(ValueError, match="gtp"):
        Engine(P.derive(P.BASE), b)
    b = c1_batch(ntraj=2)
    b.box = np.zeros(3, np.float32)
    with pytest.raises(ValueError, match="box"):
        Engine(P.derive(P.BASE), b)
