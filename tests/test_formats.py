"""`.taas` v2 container and baked ConnectomeData format (CPU only).

The reference's own tests are restated case by case:
crates/SiMuse/crates/taas-format/src/lib.rs:83-263 (round trips of every kind,
wrong kind, corruption, magic, files, TAAS1 migration, snapshot of a
"creature"), migration.rs:157-225, and pillars-connectome/src/data.rs:620-760
(receptor mappings and the baked save/load round trips).  The CRC is pinned
by the CRC-32 check value, the byte layout by hand-assembled known answers,
and the library's C++ writer/reader (dioma-core_amd/csrc/taas.hpp) against
the Python one byte for byte.
"""
import json
import os
import shutil
import struct
import subprocess
import zlib

import numpy as np
import pytest

from simuse.formats import (BAKED_MAGIC, TAAS1_MAGIC, TAAS_MAGIC, TAAS_VERSION, ConnectomeData, TaasError,
                            TaasReader, TaasWriter, elem_size, inspect_taas, migrate_bytes)

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "dioma-core_amd", "csrc")


# ---------------------------------------------------------------------------
# .taas v2 — lib.rs tests
# ---------------------------------------------------------------------------
def sample_writer():
    w = TaasWriter()
    (w.game("dioma").add_metadata("world", "meadow").f32("weights", [1.0, 2.0, 3.0]).i32("ids", [10, 20])
     .i64("big", [1 << 40]).u64("stamps", [42]).u8("mask", [0, 255, 7]).str("name", "creature-7")
     .json("genome", {"mut": 0.01, "tags": ["a"]}).blob("raw", bytes([9, 8, 7])))
    return w


def test_roundtrip_all_kinds():
    r = TaasReader.from_bytes(sample_writer().to_bytes())
    assert r.manifest.version == TAAS_VERSION
    assert r.manifest.game == "dioma"
    assert r.metadata()["world"] == "meadow"
    assert r.get_f32("weights").tolist() == [1.0, 2.0, 3.0]
    assert r.get_i32("ids").tolist() == [10, 20]
    assert r.get_i64("big").tolist() == [1 << 40]
    assert r.get_u64("stamps").tolist() == [42]
    assert r.get_u8("mask").tolist() == [0, 255, 7]
    assert r.get_str("name") == "creature-7"
    assert r.get_json("genome")["mut"] == 0.01
    assert r.get_blob("raw") == bytes([9, 8, 7])
    assert r.has_section("weights") and not r.has_section("nope")
    assert len(r.list_sections()) == 8
    r.validate()


def test_wrong_kind_rejected():
    w = TaasWriter()
    w.i32("x", [1])
    r = TaasReader.from_bytes(w.to_bytes())
    with pytest.raises(TaasError, match="is not f32"):
        r.get_f32("x")


def test_corruption_detected():
    w = TaasWriter()
    w.f32("data", np.ones(64))
    b = bytearray(w.to_bytes())
    b[-1] ^= 0xFF
    r = TaasReader.from_bytes(bytes(b))  # the manifest still parses
    with pytest.raises(TaasError, match="CRC32 mismatch"):
        r.get_f32("data")
    with pytest.raises(TaasError):
        r.validate()


def test_invalid_magic_and_truncation():
    with pytest.raises(TaasError):
        TaasReader.from_bytes(b"NOTTAAS\x00\x00\x00")
    b = sample_writer().to_bytes()
    with pytest.raises(TaasError, match="too short"):
        TaasReader.from_bytes(b[:10])
    with pytest.raises(TaasError, match="truncated"):
        TaasReader.from_bytes(b[:40])
    with pytest.raises(TaasError, match="exceeds data"):
        TaasReader.from_bytes(b[:-2])


def test_zstd_refused():
    # lib.rs:148-175 round-trips zstd sections; zstd is absent here, so both
    # directions fail loudly instead of writing or reading garbage
    with pytest.raises(TaasError, match="zstd"):
        TaasWriter().set_compression(3)
    w = TaasWriter().created("2026-01-01T00:00:00Z")
    w.f32("c", [1.0])
    man = w.manifest()
    man.sections[0].compression = "zstd"
    man.sections[0].uncompressed_size = 4
    header = man.to_json_bytes()
    b = TAAS_MAGIC + struct.pack("<Q", len(header)) + header + np.float32([1.0]).tobytes()
    r = TaasReader.from_bytes(b)
    assert r.section("c").compression == "zstd"
    with pytest.raises(TaasError, match="zstd"):
        r.get_f32("c")


def test_file_roundtrip(tmp_path):
    p = tmp_path / "x.taas"
    w = TaasWriter()
    w.f32("x", [1.5, -2.5])
    w.to_file(p)
    assert TaasReader.from_file(p).get_f32("x").tolist() == [1.5, -2.5]


def test_section_kind_mapping():
    assert elem_size("f32") == 4 and elem_size("i64") == 8 and elem_size("u8") == 1
    assert elem_size("json") is None
    with pytest.raises(TaasError):
        elem_size("nope")


def test_creature_save_restore():
    weights = (np.arange(120) * 0.01).astype(np.float32)
    w = TaasWriter()
    (w.game("dioma").u64("tick", [1234]).f32("brain_weights", weights).f32("brain_state", np.full(120, 0.5))
     .json("genome", {"grow_seed": 777, "mut": 0.05}).u8("instincts", [1, 0, 1]))
    r = TaasReader.from_bytes(w.to_bytes())
    assert r.get_u64("tick").tolist() == [1234]
    assert np.array_equal(r.get_f32("brain_weights"), weights)
    assert np.array_equal(r.get_f32("brain_state"), np.full(120, 0.5, np.float32))
    assert r.get_json("genome")["grow_seed"] == 777
    assert r.get_u8("instincts").tolist() == [1, 0, 1]


def test_inspect_output(tmp_path):
    p = tmp_path / "i.taas"
    TaasWriter().game("dioma").f32("test", [1.0, 2.0, 3.0]).add_metadata("world", "meadow").to_file(p)
    out = inspect_taas(p)
    assert "test" in out and "dioma" in out and "meadow" in out


def _taas1(components, meta):
    comps, off, data = [], 0, b""
    for name, dtype, count, raw in components:
        comps.append({"name": name, "component_type": "mig", "dtype": dtype, "shape": [count], "offset": off,
                      "size": len(raw)})
        off += len(raw)
        data += raw
    man = {"version": "0.2.0", "created": "2026-01-01T00:00:00Z", "components": comps, "metadata": meta}
    header = json.dumps(man).encode()
    return TAAS1_MAGIC + struct.pack("<Q", len(header)) + header + data


def test_migrate_taas1_roundtrip():
    old = _taas1([("weights", "f32", 4, np.float32([1, 2, 3, 4]).tobytes()),
                  ("ids", "i32", 2, np.int32([1, 2]).tobytes())], {"model": "mig-test"})
    new = migrate_bytes(old)
    assert new.startswith(TAAS_MAGIC)
    r = TaasReader.from_bytes(new)
    assert r.manifest.version == TAAS_VERSION
    assert r.get_f32("weights").tolist() == [1.0, 2.0, 3.0, 4.0]
    assert r.get_i32("ids").tolist() == [1, 2]
    assert r.metadata()["model"] == "mig-test"
    assert r.metadata()["migrated_from"] == "0.2.0"
    assert r.section("weights").crc32 is not None
    # from_bytes migrates transparently (lib.rs:177-209)
    assert TaasReader.from_bytes(old).get_f32("weights").tolist() == [1.0, 2.0, 3.0, 4.0]
    with pytest.raises(TaasError, match="Not a .taas"):
        migrate_bytes(b"XXXXX" + bytes(8))


# ---------------------------------------------------------------------------
# Byte-level known answers
# ---------------------------------------------------------------------------
def test_crc32_check_value():
    # CRC-32 (IEEE 802.3, reflected 0xEDB88320, init and xorout 0xFFFFFFFF):
    # the catalogued check value of "123456789" is 0xCBF43926
    from simuse.formats import crc32
    assert crc32(b"123456789") == 0xCBF43926
    assert crc32(b"") == 0


def test_manifest_bytes_known_answer():
    # serde_json of TaasManifest / SectionEntry (types.rs:73-100): compact,
    # declared field order, absent options skipped, metadata keys sorted
    w = TaasWriter().game("dioma").created("2026-01-01T00:00:00Z")
    w.add_metadata("zeta", "1").add_metadata("alpha", "2")
    w.f32("w", [1.0])
    b = w.to_bytes()
    crc = 0xACA16A6A  # CRC-32 of 00 00 80 3f
    assert zlib.crc32(b"\x00\x00\x80\x3f") == crc
    manifest = ('{"version":"2.0.0","created":"2026-01-01T00:00:00Z","game":"dioma","sections":'
                '[{"name":"w","kind":"f32","count":1,"offset":0,"size":4,"crc32":%d}],'
                '"metadata":{"alpha":"2","zeta":"1"}}' % crc).encode()
    assert b == b"TAAS2" + len(manifest).to_bytes(8, "little") + manifest + b"\x00\x00\x80\x3f"
    # no game: the key is absent, not null
    b2 = TaasWriter().created("t").to_bytes()
    assert b2[13:] == b'{"version":"2.0.0","created":"t","sections":[],"metadata":{}}'


# ---------------------------------------------------------------------------
# The library's C++ writer/reader against the Python one
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def taas_check(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("taas") / "taas_check")
    subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", "-I", CSRC, os.path.join(HERE, "taas_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    return exe


def _cpp(exe, *args):
    return subprocess.run([exe, *map(str, args)], capture_output=True, text=True)


def test_cpp_crc_matches(taas_check):
    for s in ["123456789", "", "creature-7", "x" * 1000]:
        assert int(_cpp(taas_check, "crc", s).stdout) == zlib.crc32(s.encode())


def test_cpp_writer_bytes_match_python(taas_check, tmp_path):
    p = tmp_path / "cpp.taas"
    assert _cpp(taas_check, "write", p).returncode == 0
    w = TaasWriter().game("dioma").created("2026-01-01T00:00:00Z")
    w.add_metadata("world", "meadow").add_metadata('a"quote', "tab\tnew\nline")
    (w.f32("weights", [1.0, 2.0, 3.0]).i32("ids", [10, 20]).i64("big", [1 << 40]).u64("stamps", [42])
     .u8("mask", [0, 255, 7]).str("name", "creature-7"))
    w.push("genome", "json", 25, b'{"mut":0.01,"tags":["a"]}')
    w.blob("raw", bytes([9, 8, 7])).f32("empty", [])
    assert p.read_bytes() == w.to_bytes()
    r = TaasReader.from_file(p)
    r.validate()
    assert r.metadata()['a"quote'] == "tab\tnew\nline"


def test_cpp_reader_reads_python_files(taas_check, tmp_path):
    p = tmp_path / "py.taas"
    w = sample_writer()
    w.add_metadata("uni", "é中")
    w.to_file(p)
    res = _cpp(taas_check, "read", p)
    assert res.returncode == 0, res.stdout
    lines = res.stdout.splitlines()
    assert lines[0] == "version 2.0.0"
    secs = [ln.split() for ln in lines[1:] if not ln.startswith("meta ")]
    assert [s[0] for s in secs] == ["weights", "ids", "big", "stamps", "mask", "name", "genome", "raw"]
    assert all(s[-1] == "ok" for s in secs)
    meta = dict(json.loads("[%s]" % ln[5:].replace('" "', '","')) for ln in lines if ln.startswith("meta "))
    assert meta == {"world": "meadow", "uni": "é中"}
    # corruption and a bad magic are reported by the C++ reader too
    b = bytearray(p.read_bytes())
    b[-1] ^= 0xFF
    p.write_bytes(bytes(b))
    res = _cpp(taas_check, "read", p)
    assert "CRC mismatch in raw" in res.stdout
    p.write_bytes(b"NOTTAAS" + bytes(10))
    res = _cpp(taas_check, "read", p)
    assert res.returncode == 1 and "invalid magic" in res.stdout


# ---------------------------------------------------------------------------
# Baked ConnectomeData — data.rs tests
# ---------------------------------------------------------------------------
def test_build_receptor_types_mammal_and_insect():
    c = ConnectomeData([0, 1, 2, 3], [1, 2, 0], [0.5, 1.0, 0.8], [1.0, -1.0, 1.0])
    c.build_receptor_types()
    assert c.receptor_types.tolist() == [0, 2, 0]
    c.build_receptor_types()  # second call is a no-op
    assert c.receptor_types.size == 3
    c2 = ConnectomeData([0, 1, 2, 3], [1, 2, 0], [0.5, 1.0, 0.8], [1.0, -1.0, 1.0])
    c2.build_receptor_types("insect")
    assert c2.receptor_types.tolist() == [4, 5, 4]


def test_build_receptor_types_flywire_and_csv():
    c = ConnectomeData([0, 1, 2, 3, 4], [1, 2, 0, 3], [0.5, 1.0, 0.8, 0.3], [1.0, -1.0, 1.0, 1.0])
    c.build_receptor_types_flywire([1.0, -1.0, 2.0, -2.0])
    assert c.receptor_types.tolist() == [4, 2, 5, 6]
    c = ConnectomeData([0, 1, 2, 3, 4, 5], [0] * 5, [1.0] * 5, [1.0] * 5)
    c.build_receptor_types_from_csv(["ACh", "GABA", "GLUT", "DA", "SER"])
    assert c.receptor_types.tolist() == [4, 2, 5, 6, 6]


def test_receptor_rows_are_spread_per_synapse():
    # rows of 2 synapses each: every synapse takes its presynaptic neuron's code
    c = ConnectomeData([0, 2, 4, 6], [1, 2, 0, 1, 2, 0], [0.5, 1.0, 0.8, 0.3, 0.2, 0.7], [1.0, -1.0, 1.0])
    c.build_receptor_types()
    assert c.receptor_types.tolist() == [0, 0, 2, 2, 0, 0]


def test_save_load_receptor_types(tmp_path):
    c = ConnectomeData([0, 2, 4, 6], [1, 2, 0, 1, 2, 0], [0.5, 1.0, 0.8, 0.3, 0.2, 0.7], [1.0, -1.0, 1.0])
    c.build_receptor_types()
    p = tmp_path / "c.taas"
    c.save_baked(p)
    d = ConnectomeData.load_baked(p)
    assert d.receptor_types.tolist() == c.receptor_types.tolist()
    assert d.n_neurons == 3 and d.n_synapses == 6
    assert np.array_equal(d.values, c.values) and np.array_equal(d.row_ptr, c.row_ptr)
    assert np.array_equal(d.col_indices, c.col_indices) and np.array_equal(d.nt_types, c.nt_types)


def test_save_load_empty_receptor_types():
    c = ConnectomeData([0, 2, 4, 6], [1, 2, 0, 1, 2, 0], [0.5] * 6, [1.0, -1.0, 1.0])
    d = ConnectomeData.from_baked_bytes(c.to_baked_bytes())
    assert d.receptor_types.size == 0
    # a file that ends before the flag byte loads the same (data.rs:364-373)
    assert ConnectomeData.from_baked_bytes(c.to_baked_bytes()[:-1]).receptor_types.size == 0


def test_baked_bytes_known_answer():
    # data.rs:292-316 on a little-endian host: magic 0x54414153 ("SAAT"),
    # version 1, n = 2, s = 1, row_ptr, col, values, nt_types, flag, receptors
    c = ConnectomeData([0, 1, 1], [1], [0.5], [1.0, -1.0], [4])
    expect = (b"SAAT" + b"\x01\x00\x00\x00" + (2).to_bytes(8, "little") + (1).to_bytes(8, "little")
              + b"\x00\x00\x00\x00\x01\x00\x00\x00\x01\x00\x00\x00"
              + b"\x01\x00\x00\x00"
              + b"\x00\x00\x00\x3f"
              + b"\x00\x00\x80\x3f\x00\x00\x80\xbf"
              + b"\x01\x04")
    assert c.to_baked_bytes() == expect
    assert BAKED_MAGIC.to_bytes(4, "little") == b"SAAT"
    with pytest.raises(ValueError, match="bad magic"):
        ConnectomeData.from_baked_bytes(b"XXXX" + expect[4:])
    with pytest.raises(ValueError, match="truncated"):
        ConnectomeData.from_baked_bytes(expect[:30])


def test_connectome_csr_wiring():
    # Brain::from_connectome (brain.rs:19-29,63-83): i32 entries clamped at 0,
    # weights = values, delays 0
    c = ConnectomeData([0, 2, 3, 3], [1, 2, 0], [0.5, -1.0, 2.0], [1.0, -1.0, 1.0])
    rp, col, w, d = c.csr()
    assert rp.dtype == np.uint32 and rp.tolist() == [0, 2, 3, 3]
    assert col.tolist() == [1, 2, 0]
    assert w.tolist() == [0.5, -1.0, 2.0]
    assert d.tolist() == [0, 0, 0]
    bad = ConnectomeData([0, 1, 1], [-5], [1.0], [1.0, 1.0])
    assert bad.csr()[1].tolist() == [0]


@pytest.mark.parametrize("manifest,ok", [
    (b'{"version":"2.0.0","created":"\\ud83d\\ude00","sections":[],"metadata":{"k":"\\u00e9"}}', True),
    (b'{"version":"2.0.0","created":"\\uZZZZ","sections":[],"metadata":{}}', False),
    (b'{"version":"2.0.0","created":"\\ud83d","sections":[],"metadata":{}}', False),
    (b'{"version":"2.0.0","created":"x","sections":[{"name":"a","kind":"f32","count":99999999999999999999999,'
     b'"offset":0,"size":4}],"metadata":{}}', False),
    (b'{"version":"2.0.0","created":"x","sections":[{"name":"a","kind":"f32","count":1,'
     b'"offset":18446744073709551615,"size":4}],"metadata":{}}', False),
    (b'{"version":"2.0.0","created":"x","sections":[{"name":"a","kind":"f32","count":"one",'
     b'"offset":0,"size":4}],"metadata":{}}', False),
    (b'{"version":"2.0.0","created":"x","sections":[', False),
    (b'[]', False),
])
def test_cpp_reader_rejects_malformed_manifests(taas_check, tmp_path, manifest, ok):
    # the library's reader reports an error for every malformed manifest
    # instead of throwing (nothing may cross the C ABI); the Python reader agrees
    p = tmp_path / "m.taas"
    p.write_bytes(TAAS_MAGIC + struct.pack("<Q", len(manifest)) + manifest + b"\x00\x00\x80\x3f")
    res = _cpp(taas_check, "read", p)
    assert res.returncode == (0 if ok else 1), (res.returncode, res.stdout, res.stderr)
    if ok:
        assert "meta \"k\" \"é\"" in res.stdout
        TaasReader.from_file(p)
    else:
        assert res.stdout.startswith("error ")
        with pytest.raises(TaasError):
            TaasReader.from_file(p)
