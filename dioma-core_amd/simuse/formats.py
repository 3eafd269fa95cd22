"""On-disk formats around the network step.

``.taas`` v2 snapshot container (crates/SiMuse/crates/taas-format/src/):

    b"TAAS2" | manifest length, u64 LE | manifest, UTF-8 JSON | payload

The manifest (types.rs:73-100) is ``{"version":"2.0.0","created":...,
["game":...,] "sections":[...], "metadata":{...}}``; each section entry is
``{"name","kind","count","offset","size"[,"crc32"][,"compression"]
[,"uncompressed_size"]}`` with kind one of f32 i32 i64 u64 u8 str json blob,
count = elements (bytes for str/json/blob), offset and size in the payload
and crc32 the CRC-32 (IEEE, zlib's) of the section's uncompressed bytes.
The JSON is compact, keys in that order and metadata keys sorted — the bytes
serde_json writes for the reference's structs — so a snapshot written here
with a fixed ``created`` is byte-identical to the reference's.  TAAS1 files
are migrated on reading (migration.rs:64-150).  zstd is not available in
this build: compressed sections are refused with :class:`TaasError`.

``ConnectomeData`` baked binary (crates/SiMuse/crates/pillars-connectome/src/
data.rs:281-383): native-endian (little-endian here) ``u32 magic
0x54414153, u32 version 1, u64 n_neurons, u64 n_synapses, i32 row_ptr[n+1],
i32 col[s], f32 values[s], f32 nt_types[n], u8 has_receptors[,
u8 receptor_types[s]]``, with the receptor mappings of data.rs:437-574 and
``Brain::from_connectome``'s wiring (brain.rs:19-29,56-83) into a network.
"""
from __future__ import annotations

import datetime as _dt
import json
import struct
import zlib
from collections import OrderedDict
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np

TAAS_MAGIC = b"TAAS2"
TAAS1_MAGIC = b"TAAS1"
TAAS_VERSION = "2.0.0"

# kind -> little-endian element dtype (None: raw bytes, count = length)
KINDS = OrderedDict([
    ("f32", np.dtype("<f4")), ("i32", np.dtype("<i4")), ("i64", np.dtype("<i8")), ("u64", np.dtype("<u8")),
    ("u8", np.dtype("u1")), ("str", None), ("json", None), ("blob", None),
])


class TaasError(ValueError):
    """A malformed, corrupt or unsupported ``.taas`` file or request."""


def elem_size(kind: str) -> Optional[int]:
    """Element size of a numeric kind, None for str/json/blob (types.rs:57-66)."""
    if kind not in KINDS:
        raise TaasError(f"unknown section kind {kind!r}")
    dt = KINDS[kind]
    return None if dt is None else dt.itemsize


def crc32(data: bytes) -> int:
    return zlib.crc32(data) & 0xFFFFFFFF


@dataclass
class SectionEntry:
    name: str
    kind: str
    count: int
    offset: int
    size: int
    crc32: Optional[int] = None
    compression: Optional[str] = None
    uncompressed_size: Optional[int] = None

    def to_json(self) -> "OrderedDict[str, object]":
        d = OrderedDict([("name", self.name), ("kind", self.kind), ("count", self.count),
                         ("offset", self.offset), ("size", self.size)])
        for k in ("crc32", "compression", "uncompressed_size"):
            v = getattr(self, k)
            if v is not None:
                d[k] = v
        return d

    @staticmethod
    def from_json(d: dict) -> "SectionEntry":
        def u64(x, bits=64):
            # serde's usize / u64 / u32: a JSON integer in range, not a bool or a float
            if isinstance(x, bool) or not isinstance(x, int) or not 0 <= x < (1 << bits):
                raise ValueError(f"{x!r} is not an unsigned {bits}-bit integer")
            return x
        try:
            crc = d.get("crc32")
            unc = d.get("uncompressed_size")
            comp = d.get("compression")
            if comp is not None and not isinstance(comp, str):
                raise ValueError("compression must be a string")
            return SectionEntry(str(d["name"]), str(d["kind"]), u64(d["count"]), u64(d["offset"]), u64(d["size"]),
                                None if crc is None else u64(crc, 32), comp, None if unc is None else u64(unc))
        except (KeyError, TypeError, ValueError) as e:
            raise TaasError(f"Invalid manifest JSON: bad section entry ({e})") from None


@dataclass
class TaasManifest:
    version: str
    created: str
    game: Optional[str]
    sections: List[SectionEntry]
    metadata: Dict[str, str]

    def to_json_bytes(self) -> bytes:
        d = OrderedDict([("version", self.version), ("created", self.created)])
        if self.game is not None:
            d["game"] = self.game
        d["sections"] = [s.to_json() for s in self.sections]
        d["metadata"] = OrderedDict(sorted(self.metadata.items()))
        return json.dumps(d, separators=(",", ":"), ensure_ascii=False).encode("utf-8")


def _now() -> str:
    return _dt.datetime.now(_dt.timezone.utc).strftime("%Y-%m-%dT%H:%M:%SZ")


def _assemble(manifest: TaasManifest, payload: bytes) -> bytes:
    """lib.rs:53-62: MAGIC + header length + JSON manifest + payload."""
    header = manifest.to_json_bytes()
    return TAAS_MAGIC + struct.pack("<Q", len(header)) + header + payload


class TaasWriter:
    """Named typed sections with CRC-32 (writer.rs:13-148)."""

    def __init__(self):
        self._sections: List[SectionEntry] = []
        self._payload = bytearray()
        self._metadata: Dict[str, str] = {}
        self._game: Optional[str] = None
        self._created: Optional[str] = None

    def game(self, game: str) -> "TaasWriter":
        self._game = str(game)
        return self

    def created(self, timestamp: str) -> "TaasWriter":
        """Fix the timestamp so that equal snapshots give equal bytes."""
        self._created = str(timestamp)
        return self

    def set_compression(self, level: int) -> "TaasWriter":
        raise TaasError("zstd compression is not available in this build")

    def disable_compression(self) -> "TaasWriter":
        return self

    def add_metadata(self, key: str, value: str) -> "TaasWriter":
        self._metadata[str(key)] = str(value)
        return self

    def push(self, name: str, kind: str, count: int, data: bytes) -> "TaasWriter":
        elem_size(kind)  # validates the kind
        data = bytes(data)
        self._sections.append(SectionEntry(str(name), kind, int(count), len(self._payload), len(data), crc32(data)))
        self._payload += data
        return self

    def _array(self, name: str, kind: str, data) -> "TaasWriter":
        a = np.ascontiguousarray(np.asarray(data), dtype=KINDS[kind])
        return self.push(name, kind, a.size, a.tobytes())

    def f32(self, name: str, data) -> "TaasWriter":
        return self._array(name, "f32", data)

    def i32(self, name: str, data) -> "TaasWriter":
        return self._array(name, "i32", data)

    def i64(self, name: str, data) -> "TaasWriter":
        return self._array(name, "i64", data)

    def u64(self, name: str, data) -> "TaasWriter":
        return self._array(name, "u64", data)

    def u8(self, name: str, data) -> "TaasWriter":
        return self._array(name, "u8", data)

    def str(self, name: str, data: str) -> "TaasWriter":
        b = data.encode("utf-8")
        return self.push(name, "str", len(b), b)

    def json(self, name: str, value) -> "TaasWriter":
        b = json.dumps(value, separators=(",", ":"), ensure_ascii=False).encode("utf-8")
        return self.push(name, "json", len(b), b)

    def blob(self, name: str, data: bytes) -> "TaasWriter":
        b = bytes(data)
        return self.push(name, "blob", len(b), b)

    def manifest(self) -> TaasManifest:
        return TaasManifest(TAAS_VERSION, self._created if self._created is not None else _now(), self._game,
                            list(self._sections), dict(self._metadata))

    def to_bytes(self) -> bytes:
        return _assemble(self.manifest(), bytes(self._payload))

    def to_file(self, path) -> None:
        with open(path, "wb") as f:
            f.write(self.to_bytes())


def _parse_header(b: bytes, magic: bytes):
    if len(b) < 13:
        raise TaasError("File too short")
    if b[:5] != magic:
        raise TaasError(f"Invalid magic: expected {magic.decode()}, got {b[:5]!r}")
    (hl,) = struct.unpack_from("<Q", b, 5)
    if 13 + hl > len(b):
        raise TaasError("Header truncated")
    try:
        man = json.loads(b[13:13 + hl].decode("utf-8"))
        # a lone surrogate escape is not a Unicode string (serde_json refuses it)
        json.dumps(man, ensure_ascii=False).encode("utf-8")
    except (UnicodeDecodeError, UnicodeEncodeError, json.JSONDecodeError, RecursionError) as e:
        raise TaasError(f"Invalid manifest JSON: {e}") from None
    if not isinstance(man, dict):
        raise TaasError("Invalid manifest JSON: not an object")
    return man, 13 + hl


def migrate_bytes(b: bytes) -> bytes:
    """TAAS2 unchanged; TAAS1 components rebuilt as v2 sections (migration.rs:64-150)."""
    b = bytes(b)
    if b.startswith(TAAS_MAGIC):
        return b
    if not b.startswith(TAAS1_MAGIC):
        raise TaasError(f"Not a .taas file: magic {b[:5]!r}")
    man, pos = _parse_header(b, TAAS1_MAGIC)
    data = b[pos:]
    w = TaasWriter()
    w.add_metadata("migrated_from", str(man.get("version", "")))
    for k, v in (man.get("metadata") or {}).items():
        w.add_metadata(k, v)
    for comp in man.get("components") or []:
        name, off, size = comp["name"], int(comp["offset"]), int(comp["size"])
        if off + size > len(data):
            raise TaasError(f"TAAS1 component '{name}' (offset={off}, size={size}) out of bounds")
        raw = data[off:off + size]
        if comp.get("compression") == "zstd":
            raise TaasError(f"TAAS1 '{name}': zstd-compressed components are not supported")
        if comp.get("crc32") is not None and crc32(raw) != int(comp["crc32"]):
            raise TaasError(f"TAAS1 component '{name}' CRC32 mismatch")
        kind = comp.get("dtype")
        if kind not in KINDS:
            raise TaasError(f"TAAS1 component '{name}' has unknown dtype '{kind}'")
        es = elem_size(kind)
        w.push(name, kind, len(raw) // es if es else len(raw), raw)
    return w.to_bytes()


class TaasReader:
    """Sections by name with kind, size and CRC checks (reader.rs:25-281)."""

    def __init__(self, manifest: TaasManifest, data: bytes, data_offset: int):
        self.manifest = manifest
        self._data = data
        self._off = data_offset

    @classmethod
    def from_bytes(cls, b: bytes) -> "TaasReader":
        b = bytes(b)
        if b.startswith(TAAS1_MAGIC):
            b = migrate_bytes(b)
        man, pos = _parse_header(b, TAAS_MAGIC)
        try:
            sections = [SectionEntry.from_json(s) for s in man["sections"]]
            meta = man.get("metadata") or {}
            manifest = TaasManifest(str(man["version"]), str(man.get("created", "")), man.get("game"), sections,
                                    {str(k): str(v) for k, v in meta.items()})
        except (KeyError, TypeError, AttributeError) as e:
            raise TaasError(f"Invalid manifest JSON: {e}") from None
        for s in sections:
            if pos + s.offset + s.size > len(b):
                raise TaasError(f"Section '{s.name}' (offset={s.offset}, size={s.size}) exceeds data "
                                f"(len={len(b) - pos})")
        return cls(manifest, b, pos)

    @classmethod
    def from_file(cls, path) -> "TaasReader":
        with open(path, "rb") as f:
            return cls.from_bytes(f.read())

    # -- sections ---------------------------------------------------------------
    def section(self, name: str) -> SectionEntry:
        for s in self.manifest.sections:
            if s.name == name:
                return s
        raise TaasError(f"Section '{name}' not found in manifest")

    def has_section(self, name: str) -> bool:
        return any(s.name == name for s in self.manifest.sections)

    def list_sections(self) -> List[SectionEntry]:
        return list(self.manifest.sections)

    def metadata(self) -> Dict[str, str]:
        return self.manifest.metadata

    def total_size(self) -> int:
        return len(self._data)

    def _read(self, s: SectionEntry) -> bytes:
        raw = self._data[self._off + s.offset:self._off + s.offset + s.size]
        if s.compression is not None:
            raise TaasError(f"Section '{s.name}' is {s.compression}-compressed; zstd is not available in this build")
        if s.crc32 is not None and crc32(raw) != s.crc32:
            raise TaasError(f"Section '{s.name}' CRC32 mismatch: expected {s.crc32:08x}, got {crc32(raw):08x}")
        return raw

    def _typed(self, name: str, kind: str) -> np.ndarray:
        s = self.section(name)
        if s.kind != kind:
            raise TaasError(f"Section '{name}' is not {kind}, got {s.kind}")
        raw = self._read(s)
        es = elem_size(kind)
        if len(raw) != s.count * es:
            raise TaasError(f"Section '{name}' {kind}: data len {len(raw)} != count {s.count} * {es}")
        return np.frombuffer(raw, dtype=KINDS[kind]).copy()

    def get_f32(self, name: str) -> np.ndarray:
        return self._typed(name, "f32")

    def get_i32(self, name: str) -> np.ndarray:
        return self._typed(name, "i32")

    def get_i64(self, name: str) -> np.ndarray:
        return self._typed(name, "i64")

    def get_u64(self, name: str) -> np.ndarray:
        return self._typed(name, "u64")

    def get_u8(self, name: str) -> np.ndarray:
        return self._typed(name, "u8")

    def get_blob(self, name: str) -> bytes:
        return self._read(self.section(name))

    def get_str(self, name: str) -> str:
        s = self.section(name)
        if s.kind != "str":
            raise TaasError(f"Section '{name}' is not str, got {s.kind}")
        try:
            return self._read(s).decode("utf-8")
        except UnicodeDecodeError as e:
            raise TaasError(f"Section '{name}' is not valid UTF-8: {e}") from None

    def get_json(self, name: str):
        s = self.section(name)
        if s.kind != "json":
            raise TaasError(f"Section '{name}' is not json, got {s.kind}")
        try:
            return json.loads(self._read(s).decode("utf-8"))
        except (UnicodeDecodeError, json.JSONDecodeError) as e:
            raise TaasError(f"Section '{name}' bad JSON: {e}") from None

    def validate(self) -> None:
        """Read every section and check its CRC."""
        for s in self.manifest.sections:
            self._read(s)


def inspect_taas(path) -> str:
    """Human-readable summary of a file (inspect.rs)."""
    r = TaasReader.from_file(path)
    m = r.manifest
    lines = [f"TAAS v{m.version}  created {m.created}" + (f"  game {m.game}" if m.game else ""),
             f"{len(m.sections)} sections, {r.total_size()} bytes"]
    for s in m.sections:
        lines.append(f"  {s.name:<24} {s.kind:<5} count={s.count:<10} size={s.size:<10}"
                     + (f" crc32={s.crc32:08x}" if s.crc32 is not None else ""))
    for k, v in sorted(m.metadata.items()):
        lines.append(f"  meta {k} = {v}")
    return "\n".join(lines)


# ---------------------------------------------------------------------------
# Baked ConnectomeData
# ---------------------------------------------------------------------------
BAKED_MAGIC = 0x54414153
BAKED_VERSION = 1

# receptor codes (0 AMPA, 2 GABA_A, 4 nAChR, 5 GluCl, 6 mAChR)
_CSV_RECEPTOR = {"ACh": 4, "GABA": 2, "GLUT": 5, "DA": 6, "DOPAMINE": 6, "SER": 6, "SEROTONIN": 6, "5HT": 6,
                 "OCT": 6, "OCTOPAMINE": 6}


def _per_synapse(row_ptr: np.ndarray, n_syn: int, rec_of_pre) -> np.ndarray:
    """Spread one receptor code per presynaptic row over its synapses, in
    synapse order, as data.rs:443-462 does (length max(n_syn, 1))."""
    out = np.zeros(max(n_syn, 1), dtype=np.uint8)
    n = len(row_ptr) - 1 if len(row_ptr) > 1 else 0
    k = 0
    for pre in range(n):
        start = max(int(row_ptr[pre]), 0)
        e = int(row_ptr[pre + 1])
        end = n_syn if e < 0 else min(e, n_syn)  # a negative i32 read as usize exceeds n_syn
        cnt = max(end - start, 0)
        lo, hi = min(k, out.size), min(k + cnt, out.size)
        out[lo:hi] = rec_of_pre(pre)
        k += cnt
    return out


class ConnectomeData:
    """CSR connectome with per-neuron NT signs and optional per-synapse
    receptor codes (data.rs:61-103)."""

    def __init__(self, row_ptr, col_indices, values, nt_types, receptor_types=None):
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
        self.col_indices = np.ascontiguousarray(col_indices, dtype=np.int32)
        self.values = np.ascontiguousarray(values, dtype=np.float32)
        self.nt_types = np.ascontiguousarray(nt_types, dtype=np.float32)
        self.receptor_types = np.ascontiguousarray(receptor_types if receptor_types is not None else [],
                                                   dtype=np.uint8)
        self.n_neurons = self.row_ptr.size - 1 if self.row_ptr.size > 1 else 0
        self.n_synapses = self.values.size
        if self.col_indices.size != self.n_synapses:
            raise ValueError("col_indices length must match values length")
        if self.nt_types.size and self.nt_types.size < self.n_neurons:
            raise ValueError("nt_types must have at least n_neurons elements")

    # -- receptor types -----------------------------------------------------------
    def build_receptor_types(self, phenotype: str = "mammal") -> None:
        """exc -> AMPA (0) / nAChR (4), inh -> GABA_A (2) / GluCl (5) by the
        sign of the presynaptic NT value (data.rs:44-58,437-463)."""
        if self.receptor_types.size:
            return
        exc, inh = {"mammal": (0, 2), "insect": (4, 5)}[phenotype]

        def rec(pre):
            nt = float(self.nt_types[pre]) if pre < self.nt_types.size else 1.0
            return exc if nt > 0.0 else inh
        self.receptor_types = _per_synapse(self.row_ptr, self.n_synapses, rec)

    def build_receptor_types_flywire(self, nt_map: Sequence[float]) -> None:
        """FlyWire mapping (data.rs:477-521): +2 GluCl (5), +1 nAChR (4),
        -2 mAChR (6), otherwise GABA_A (2)."""
        if self.receptor_types.size:
            return
        nt_map = np.asarray(nt_map, dtype=np.float32)

        def rec(pre):
            nt = float(nt_map[pre]) if pre < nt_map.size else 1.0
            return 5 if nt > 1.5 else 4 if nt > 0.5 else 6 if nt < -1.5 else 2
        self.receptor_types = _per_synapse(self.row_ptr, self.n_synapses, rec)

    def build_receptor_types_from_csv(self, nt_labels: Sequence[str]) -> None:
        """Textual NT labels (data.rs:538-574); unknown labels give AMPA (0)."""
        if self.receptor_types.size:
            return

        def rec(pre):
            return _CSV_RECEPTOR.get(nt_labels[pre], 0) if pre < len(nt_labels) else 0
        self.receptor_types = _per_synapse(self.row_ptr, self.n_synapses, rec)

    # -- baked binary ---------------------------------------------------------------
    def to_baked_bytes(self) -> bytes:
        """save_baked (data.rs:292-316)."""
        parts = [struct.pack("<IIQQ", BAKED_MAGIC, BAKED_VERSION, self.n_neurons, self.n_synapses),
                 self.row_ptr.astype("<i4").tobytes(), self.col_indices.astype("<i4").tobytes(),
                 self.values.astype("<f4").tobytes(), self.nt_types.astype("<f4").tobytes()]
        if self.receptor_types.size:
            parts += [b"\x01", self.receptor_types.tobytes()]
        else:
            parts.append(b"\x00")
        return b"".join(parts)

    def save_baked(self, path) -> None:
        with open(path, "wb") as f:
            f.write(self.to_baked_bytes())

    @classmethod
    def from_baked_bytes(cls, b: bytes) -> "ConnectomeData":
        """load_baked (data.rs:321-383): magic checked, version read and
        ignored, nt_types read as n_neurons values, receptor types only when
        the flag byte is 1."""
        b = bytes(b)
        if len(b) < 24:
            raise ValueError(".taas: file too short")
        magic, _version, n, s = struct.unpack_from("<IIQQ", b, 0)
        if magic != BAKED_MAGIC:
            raise ValueError(f".taas: bad magic {magic:#x}, expected {BAKED_MAGIC:#x}")
        pos = 24
        need = 4 * (n + 1) + 4 * s + 4 * s + 4 * n
        if pos + need > len(b):
            raise ValueError(".taas: truncated arrays")

        def take(count, dt):
            nonlocal pos
            a = np.frombuffer(b, dtype=dt, count=count, offset=pos).copy()
            pos += a.nbytes
            return a
        row_ptr = take(n + 1, "<i4")
        col = take(s, "<i4")
        values = take(s, "<f4")
        nt = take(n, "<f4")
        rec = np.zeros(0, dtype=np.uint8)
        if pos < len(b) and b[pos] == 1:
            pos += 1
            if pos + s > len(b):
                raise ValueError(".taas: truncated receptor types")
            rec = take(s, "u1")
        c = cls.__new__(cls)
        c.row_ptr, c.col_indices, c.values, c.nt_types, c.receptor_types = row_ptr, col, values, nt, rec
        c.n_neurons, c.n_synapses = int(n), int(s)
        return c

    @classmethod
    def load_baked(cls, path) -> "ConnectomeData":
        with open(path, "rb") as f:
            return cls.from_baked_bytes(f.read())

    # -- wiring ------------------------------------------------------------------
    def csr(self):
        """``Brain::from_connectome``'s graph (brain.rs:19-29,63-83): the i32
        CSR clamped at 0 and read as u32, weights = values, every delay 0."""
        rp = np.maximum(self.row_ptr.astype(np.int64), 0).astype(np.uint32)
        col = np.maximum(self.col_indices.astype(np.int64), 0).astype(np.uint32)
        return rp, col, self.values.copy(), np.zeros(self.n_synapses, dtype=np.uint8)

    def to_network(self, pop):
        """One population ``pop`` (a NetworkPop of n_neurons neurons) wired as
        the connectome says; the library validates the CSR."""
        from .network import CompiledNetwork
        if pop.n != self.n_neurons:
            raise ValueError(f"population has {pop.n} neurons, connectome {self.n_neurons}")
        rp, col, w, d = self.csr()
        return CompiledNetwork([pop], rp, col, w, d)
