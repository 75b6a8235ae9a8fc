"""The C-ABI library loads and exports every function declared in include/*.h (no GPU)."""
import os
import re

import pgs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(pgs_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_all_declared_symbols():
    L = pgs.lib()
    names = declared_functions()
    assert names, "no declarations found"
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(pgs.EXPORTED_SYMBOLS) >= names


def test_stream_generator_is_deterministic():
    a = pgs.gen_transfers(0x5EED0002, 12, 3, 64)
    b = pgs.gen_transfers(0x5EED0002, 12, 3, 64)
    c = pgs.gen_transfers(0x5EED0003, 12, 3, 64)
    assert a.tobytes() == b.tobytes()
    assert a.tobytes() != c.tobytes()
    assert len(a) % 16 == 0
