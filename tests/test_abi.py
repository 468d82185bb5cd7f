"""C ABI checks that need no GPU: the built library exports every symbol that
include/mojovec_gpu.h declares, the Python binding covers them, and argument
validation (which runs before any device call) reports the reference's
messages."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mojovec_gpu.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mvg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_expected_families():
    syms = declared_symbols()
    for fam in ("mvg_flat_", "mvg_sq8_", "mvg_kmeans_", "mvg_pq_", "mvg_ivfpq_", "mvg_ivfflat_",
                "mvg_coll_"):
        assert any(s.startswith(fam) for s in syms), fam


def test_library_exports_every_declared_symbol():
    import ctypes

    from mojovec_amd import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert set(declared_symbols()) <= set(_lib.SIGNATURES), set(declared_symbols()) - set(_lib.SIGNATURES)


def test_package_imports_and_reports_version():
    import mojovec_amd

    assert mojovec_amd._lib.load().mvg_version().startswith(b"mojovec-gpu")


@pytest.mark.parametrize(
    "args,match",
    [
        ((0,), "dimension"),
        ((10, 8, 3), "divide dimension"),
        ((8, 4, 2, 5), "nprobe"),
        ((8, 0, 2), "nlist"),
        ((8, 4, 2, None, "hamming"), "metric"),
    ],
)
def test_collection_validation_messages(args, match):
    import mojovec_amd

    with pytest.raises(mojovec_amd.MojoVecError, match=match):
        mojovec_amd.IVFPQCollection(*args)


def test_index_validation_messages():
    import mojovec_amd

    with pytest.raises(mojovec_amd.MojoVecError, match="nlist must be between 1 and 1000000"):
        mojovec_amd.IndexIVFPQ(8, 0, 2)
    with pytest.raises(mojovec_amd.MojoVecError, match="M must divide dimension"):
        mojovec_amd.IndexIVFPQ(10, 4, 3)
    with pytest.raises(mojovec_amd.MojoVecError, match="PQ M must divide dimension"):
        mojovec_amd.ProductQuantizer(10, 3)
    with pytest.raises(mojovec_amd.MojoVecError, match="IVF-Flat dimension must be positive"):
        mojovec_amd.IndexIVFFlat(0, 4)
    with pytest.raises(mojovec_amd.MojoVecError, match="IVF-Flat nlist must be between 1 and 1000000"):
        mojovec_amd.IndexIVFFlat(8, 0)
    with pytest.raises(mojovec_amd.MojoVecError, match="metric"):
        mojovec_amd.IndexIVFFlat(8, 4, 7)


def test_null_handles_are_rejected_before_any_device_work():
    import ctypes

    import mojovec_amd

    lib = mojovec_amd._lib.load()
    out = ctypes.c_void_p()
    for create in (lib.mvg_server_create, lib.mvg_server_create_sq8):
        assert create(None, 2, 64, 1 << 20, ctypes.byref(out)) != 0
        assert b"server needs an index" in lib.mvg_last_error()
    assert lib.mvg_ivfpq_set_prefilter(None, 8) != 0
    assert lib.mvg_ivfpq_set_table_bits(None, 8) != 0
    assert lib.mvg_ivfpq_set_tile_min_batch(None, 1) != 0
    assert lib.mvg_last_error()
    # the later entry points check their handle first as well
    assert lib.mvg_ivfpq_add_with_ids(None, None, 1, None) != 0
    assert lib.mvg_flat_get_vectors(None, None, 0, None) != 0
    assert lib.mvg_sq8_get_vector(None, 0, None) != 0
    assert lib.mvg_ivfpq_search_device_k(None, None, 1, None, 1, None, 1, 1, None, 0, None, None,
                                         None) != 0
    assert lib.mvg_ivfpq_capture_k(None, None, 1, None, 1, None, 1, 1, None, None,
                                   ctypes.byref(out)) != 0
    assert lib.mvg_graph_launch(None, None) != 0
    # PQ decode validates its shape before touching a device; n = 0 is a no-op
    assert lib.mvg_pq_decode(None, 0, 8, 2, None, None) == 0
    assert lib.mvg_pq_decode(None, 4, 8, 3, None, None) != 0
    assert b"PQ M must divide dimension" in lib.mvg_last_error()


def test_no_oracle_in_product_package():
    pkg = os.path.join(ROOT, "mojovec_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                text = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "oracle" not in text.lower() or f == "build.py", f
                assert "mvref" not in text, f
