"""Hall basis metadata on disk: rpp_hip_lie_basis_save / rpp_hip_lie_basis_load
and the RPP_HIP_LIE_CACHE_DIR cache.  Every load that installs metadata runs in
a fresh interpreter, whose process-wide metadata starts empty; the host code
only is exercised (no HIP call)."""
import json
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import rpp_hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "rough-path-primitives_amd")

HEADER = 8 + 8 + 32  # tag, (width, depth), four int64 sizes


def fnv1a(data: bytes) -> int:
    h = 1469598103934665603
    for b in data:
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def reseal(body: bytes) -> bytes:
    """The file body with a fresh checksum (a valid file with chosen contents)."""
    return body + struct.pack("<Q", fnv1a(body))


def layout(data: bytes):
    """Byte offsets of the sections of a metadata file."""
    w, d = struct.unpack_from("<ii", data, 8)
    ts, dim, l2t_nnz, t2l_nnz = struct.unpack_from("<qqqq", data, 16)
    pairs = HEADER
    db = pairs + 16 * ts
    l2t_v = db + 8 * (d + 2)
    l2t_i = l2t_v + 8 * l2t_nnz
    l2t_o = l2t_i + 8 * l2t_nnz
    t2l_v = l2t_o + 8 * ts
    t2l_i = t2l_v + 8 * t2l_nnz
    return dict(width=w, depth=d, ts=ts, dim=dim, pairs=pairs, t2l_values=t2l_v, t2l_indices=t2l_i,
                l2t_indices=l2t_i, t2l_nnz=t2l_nnz)


def child(code: str, env_extra=None) -> dict:
    env = dict(os.environ)
    env.update(env_extra or {})
    prog = f"import sys, json\nsys.path.insert(0, {PKG!r})\nimport rpp_hip, numpy as np\n" + code
    r = subprocess.run([sys.executable, "-c", prog], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


SUMMARY = """
lb = rpp_hip.lie_basis({w}, {d})
print(json.dumps(dict(pairs=lb["pairs"].tolist(), db=lb["degree_begin"].tolist(),
                      l2t=[lb["l2t"][k].tolist() for k in ("data", "indices", "offsets")],
                      t2l=[lb["t2l"][k].tolist() for k in ("data", "indices", "offsets")])))
"""


def summary_here(w, d):
    lb = rpp_hip.lie_basis(w, d)
    return dict(pairs=lb["pairs"].tolist(), db=lb["degree_begin"].tolist(),
                l2t=[lb["l2t"][k].tolist() for k in ("data", "indices", "offsets")],
                t2l=[lb["t2l"][k].tolist() for k in ("data", "indices", "offsets")])


def test_round_trip_in_a_fresh_process(tmp_path):
    path = tmp_path / "w3d5.bin"
    rpp_hip.lie_basis_save(3, 5, path)
    got = child(f"print(json.dumps(rpp_hip.lie_basis_load({str(path)!r})))\n" + SUMMARY.format(w=3, d=5))
    assert got == summary_here(3, 5)


def test_the_loaded_file_is_what_the_process_uses(tmp_path):
    """A valid file with one T2L value changed: the process that loads it
    reports the changed value, so conversions use the file, not a rebuild;
    loading the original afterwards is refused as a conflicting second
    metadata for the same shape."""
    path, edited = tmp_path / "orig.bin", tmp_path / "edited.bin"
    rpp_hip.lie_basis_save(3, 5, path)
    data = path.read_bytes()
    lo = layout(data)
    body = bytearray(data[:-8])
    struct.pack_into("<d", body, lo["t2l_values"], 12345.0)
    edited.write_bytes(reseal(bytes(body)))
    code = (f"print(rpp_hip.lie_basis_load({str(edited)!r}))\n"
            "lb = rpp_hip.lie_basis(3, 5)\n"
            "code = None\n"
            "try:\n"
            f"    rpp_hip.lie_basis_load({str(path)!r})\n"
            "except rpp_hip.RppHipError as e:\n"
            "    code = e.code\n"
            "print(json.dumps(dict(first=float(lb['t2l']['data'][0]), conflict=code)))\n")
    got = child(code)
    assert got == {"first": 12345.0, "conflict": 6}


def test_damaged_or_inconsistent_files_are_refused(tmp_path):
    good = tmp_path / "good.bin"
    rpp_hip.lie_basis_save(2, 6, good)
    data = good.read_bytes()
    lo = layout(data)
    cases = {}
    flipped = bytearray(data)
    flipped[len(data) // 2] ^= 0x40
    cases["flipped byte"] = bytes(flipped)
    cases["truncated"] = data[: len(data) // 3]
    cases["empty"] = b""
    body = bytearray(data[:-8])
    body[0:8] = b"NOTALIE!"
    cases["foreign tag"] = reseal(bytes(body))
    body = bytearray(data[:-8])
    struct.pack_into("<q", body, lo["t2l_indices"], lo["ts"] + 5)  # a T2L row beyond the Lie keys
    cases["index out of range"] = reseal(bytes(body))
    body = bytearray(data[:-8])
    struct.pack_into("<q", body, 16 + 8, lo["dim"] + 1)  # tensor dimension of another shape
    cases["wrong dimension"] = reseal(bytes(body))
    cases["trailing bytes"] = reseal(data[:-8] + b"\0" * 8)
    body = bytearray(data[:-8])
    # the last T2L entry belongs to a top-degree word: point it at a degree-1 key
    struct.pack_into("<q", body, lo["t2l_indices"] + 8 * (lo["t2l_nnz"] - 1), 0)
    cases["entry across degrees"] = reseal(bytes(body))
    # counts beyond the file (2 * true_size would overflow 64 bits) and a
    # width whose powers overflow
    body = bytearray(data[:-8])
    struct.pack_into("<q", body, 16, (1 << 62) + 3)
    cases["huge true_size"] = reseal(bytes(body))
    body = bytearray(data[:-8])
    struct.pack_into("<i", body, 8, (1 << 31) - 1)
    cases["huge width"] = reseal(bytes(body))
    for name, blob in cases.items():
        p = tmp_path / (name.replace(" ", "_") + ".bin")
        p.write_bytes(blob)
        with pytest.raises(rpp_hip.RppHipError) as e:
            rpp_hip.lie_basis_load(p)
        assert e.value.code == 6, name
    with pytest.raises(rpp_hip.RppHipError) as e:
        rpp_hip.lie_basis_load(tmp_path / "missing.bin")
    assert e.value.code == 3
    # the good file loads here: it equals the metadata this process built
    assert rpp_hip.lie_basis_load(good) == (2, 6)


def test_cache_directory_is_written_read_and_repaired(tmp_path):
    env = {"RPP_HIP_LIE_CACHE_DIR": str(tmp_path)}
    first = child(SUMMARY.format(w=2, d=6), env)
    f = tmp_path / "rpp_lie_w2_d6.bin"
    assert f.exists()
    assert first == summary_here(2, 6)
    assert not [p for p in os.listdir(tmp_path) if ".tmp." in p]
    # a valid edited cache file is what the next process uses
    data = f.read_bytes()
    lo = layout(data)
    body = bytearray(data[:-8])
    struct.pack_into("<d", body, lo["t2l_values"], -7.5)
    f.write_bytes(reseal(bytes(body)))
    second = child(SUMMARY.format(w=2, d=6), env)
    assert second["t2l"][0][0] == -7.5
    # a damaged one is rebuilt and rewritten
    damaged = bytearray(f.read_bytes())
    damaged[100] ^= 1
    f.write_bytes(bytes(damaged))
    third = child(SUMMARY.format(w=2, d=6), env)
    assert third == summary_here(2, 6)
    assert rpp_hip.lie_basis_load(f) == (2, 6)
