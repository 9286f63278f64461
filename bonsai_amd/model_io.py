"""Model file v7 (the reference's `io::save_booster` / `io::load_booster`,
src/io/model.cpp:160-387): one msgpack document

    {"bin_mappers": [{"cuts": [...], "name": ...}, ...], "config": {...},
     "init_score": f | "init_scores": [f, ...], "magic": "bonsai01",
     "trees": [{"covers", "depth", "gains", "n_leaves", "nodes": [{"default_left",
                "feature_id", "left", "right", "threshold_or_value"}, ...]}, ...],
     "version": 7}

written by nlohmann::json::to_msgpack, whose choices this writer follows so that a
loaded model saves back to the same bytes (the reference's own round-trip test,
tests/unit/test_model_io.cpp:270-297):
- object keys in byte order (std::map);
- an integer in the smallest form: positive fixint, uint8/16/32/64 for values >= 0,
  negative fixint or int8/16/32/64 below;
- a float as float32 when it lies within +-FLT_MAX and converts exactly, else float64
  (the +inf cut sentinel is float64);
- strings as fixstr/str8/str16/str32, arrays and maps in their fix/16/32 forms.

Trees are DenseTree records in threshold space (include/bonsai/tree.hpp:24-79): a leaf
has feature_id = 0xFFFFFFFF and its value in threshold_or_value; a split sends
x <= threshold (and, for NaN, default_left) to `left`.  `Model.bin_space()` turns them
into the device predictor's bin-space nodes: threshold = cuts[f][b] (right-inclusive
bins) gives split bin b, and the missing bin (the last) follows default_left.
Oblivious (levelwise) trees are kept as their raw records.
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field

import numpy as np

MAGIC = "bonsai01"
FORMAT_VERSION = 7
LEAF_FLAG = 0xFFFFFFFF
_FLT_MAX = float(np.finfo(np.float32).max)


class ModelFormatError(ValueError):
    """Not a v7 model file (bad magic, other version, truncated or malformed)."""


# ---------------------------------------------------------------------------------
# msgpack codec
# ---------------------------------------------------------------------------------

def unpack(data: bytes):
    """Decode one msgpack document (maps become dicts in file order)."""
    pos = 0

    def need(n):
        if pos + n > len(data):
            raise ModelFormatError("truncated msgpack document")

    def take(n):
        nonlocal pos
        need(n)
        b = data[pos:pos + n]
        pos += n
        return b

    def value():
        nonlocal pos
        t = take(1)[0]
        if t <= 0x7F:
            return t
        if t >= 0xE0:
            return t - 0x100
        if 0x80 <= t <= 0x8F:
            return mapping(t & 0x0F)
        if 0x90 <= t <= 0x9F:
            return [value() for _ in range(t & 0x0F)]
        if 0xA0 <= t <= 0xBF:
            return take(t & 0x1F).decode("utf-8")
        if t == 0xC0:
            return None
        if t == 0xC2:
            return False
        if t == 0xC3:
            return True
        if t == 0xCA:
            return struct.unpack(">f", take(4))[0]
        if t == 0xCB:
            return struct.unpack(">d", take(8))[0]
        fmt = {0xCC: ">B", 0xCD: ">H", 0xCE: ">I", 0xCF: ">Q", 0xD0: ">b", 0xD1: ">h", 0xD2: ">i", 0xD3: ">q"}
        if t in fmt:
            return struct.unpack(fmt[t], take(struct.calcsize(fmt[t])))[0]
        if t in (0xD9, 0xDA, 0xDB):
            n = struct.unpack({0xD9: ">B", 0xDA: ">H", 0xDB: ">I"}[t], take(1 << (t - 0xD9)))[0]
            return take(n).decode("utf-8")
        if t in (0xDC, 0xDD):
            n = struct.unpack(">H" if t == 0xDC else ">I", take(2 if t == 0xDC else 4))[0]
            return [value() for _ in range(n)]
        if t in (0xDE, 0xDF):
            n = struct.unpack(">H" if t == 0xDE else ">I", take(2 if t == 0xDE else 4))[0]
            return mapping(n)
        raise ModelFormatError(f"msgpack type 0x{t:02x} is not used by model files")

    def mapping(n):
        out = {}
        for _ in range(n):
            k = value()
            if not isinstance(k, str):
                raise ModelFormatError("model file map keys are strings")
            out[k] = value()
        return out

    v = value()
    if pos != len(data):
        raise ModelFormatError("trailing bytes after the msgpack document")
    return v


def _pack_int(v: int, out: bytearray):
    if v >= 0:
        if v <= 0x7F:
            out.append(v)
        elif v <= 0xFF:
            out += b"\xcc" + struct.pack(">B", v)
        elif v <= 0xFFFF:
            out += b"\xcd" + struct.pack(">H", v)
        elif v <= 0xFFFFFFFF:
            out += b"\xce" + struct.pack(">I", v)
        else:
            out += b"\xcf" + struct.pack(">Q", v)
    elif v >= -32:
        out.append(v + 0x100)
    elif v >= -128:
        out += b"\xd0" + struct.pack(">b", v)
    elif v >= -32768:
        out += b"\xd1" + struct.pack(">h", v)
    elif v >= -(1 << 31):
        out += b"\xd2" + struct.pack(">i", v)
    else:
        out += b"\xd3" + struct.pack(">q", v)


def _pack_len(n: int, fix: int, fix_max: int, codes: tuple, out: bytearray):
    if n <= fix_max:
        out.append(fix | n)
    elif n <= 0xFFFF and codes[0] is not None:
        out += bytes([codes[0]]) + struct.pack(">H", n)
    else:
        out += bytes([codes[1]]) + struct.pack(">I", n)


def pack(v) -> bytes:
    """Encode with nlohmann::json::to_msgpack's choices (see the module docstring)."""
    out = bytearray()

    def rec(x):
        nonlocal out
        if x is None:
            out.append(0xC0)
        elif isinstance(x, (bool, np.bool_)):
            out.append(0xC3 if x else 0xC2)
        elif isinstance(x, (int, np.integer)):
            _pack_int(int(x), out)
        elif isinstance(x, (float, np.floating)):
            d = float(x)
            f32 = float(np.float32(d)) if not math.isnan(d) else None
            if -_FLT_MAX <= d <= _FLT_MAX and f32 == d:
                out += b"\xca" + struct.pack(">f", d)
            else:
                out += b"\xcb" + struct.pack(">d", d)
        elif isinstance(x, str):
            b = x.encode("utf-8")
            if len(b) <= 31:
                out.append(0xA0 | len(b))
            elif len(b) <= 0xFF:
                out += b"\xd9" + struct.pack(">B", len(b))
            elif len(b) <= 0xFFFF:
                out += b"\xda" + struct.pack(">H", len(b))
            else:
                out += b"\xdb" + struct.pack(">I", len(b))
            out += b
        elif isinstance(x, (list, tuple)):
            _pack_len(len(x), 0x90, 15, (0xDC, 0xDD), out)
            for e in x:
                rec(e)
        elif isinstance(x, dict):
            _pack_len(len(x), 0x80, 15, (0xDE, 0xDF), out)
            for k in sorted(x, key=lambda s: s.encode("utf-8")):
                rec(k)
                rec(x[k])
        else:
            raise TypeError(f"cannot encode {type(x).__name__} in a model file")

    rec(v)
    return bytes(out)


# ---------------------------------------------------------------------------------
# Typed model
# ---------------------------------------------------------------------------------

@dataclass
class DenseTree:
    """DenseTree (include/bonsai/tree.hpp): nodes as parallel arrays plus gains/covers."""
    feature_id: np.ndarray      # u32, LEAF_FLAG for leaves
    value: np.ndarray           # f32 threshold (split) or leaf value
    left: np.ndarray            # u32
    right: np.ndarray           # u32
    default_left: np.ndarray    # bool
    gains: np.ndarray           # f32 per node
    covers: np.ndarray          # f32 per node
    depth: int
    n_leaves: int

    @staticmethod
    def from_record(r: dict) -> "DenseTree":
        nodes = r["nodes"]
        return DenseTree(
            np.array([n["feature_id"] for n in nodes], np.uint32),
            np.array([n["threshold_or_value"] for n in nodes], np.float32),
            np.array([n["left"] for n in nodes], np.uint32),
            np.array([n["right"] for n in nodes], np.uint32),
            np.array([n["default_left"] for n in nodes], bool),
            np.array(r["gains"], np.float32), np.array(r["covers"], np.float32),
            int(r["depth"]), int(r["n_leaves"]))

    def record(self) -> dict:
        nodes = [{"feature_id": int(f), "threshold_or_value": float(v), "left": int(lo), "right": int(hi),
                  "default_left": bool(d)}
                 for f, v, lo, hi, d in zip(self.feature_id, self.value, self.left, self.right, self.default_left)]
        return {"depth": self.depth, "n_leaves": self.n_leaves, "nodes": nodes,
                "gains": [float(g) for g in self.gains], "covers": [float(c) for c in self.covers]}


@dataclass
class Model:
    """A loaded v7 model: config, per-feature cuts and names, init score(s), trees."""
    config: dict
    feature_names: list
    cuts: list                  # np.float32 arrays, FLT_MAX closer + +inf sentinel included
    init_score: float | None = None
    init_scores: list | None = None
    trees: list = field(default_factory=list)   # DenseTree, or raw dicts (oblivious trees)

    @property
    def n_classes(self) -> int:
        return len(self.init_scores) if self.init_scores is not None else 1

    @property
    def learning_rate(self) -> float:
        return float(self.config["booster_config"]["learning_rate"])

    def n_bins(self) -> np.ndarray:
        return np.array([len(c) for c in self.cuts], np.uint16)

    def document(self) -> dict:
        doc = {"magic": MAGIC, "version": FORMAT_VERSION, "config": self.config,
               "bin_mappers": [{"name": n, "cuts": [float(c) for c in cs]}
                               for n, cs in zip(self.feature_names, self.cuts)],
               "trees": [t.record() if isinstance(t, DenseTree) else t for t in self.trees]}
        if self.init_scores is not None:
            doc["init_scores"] = [float(np.float32(s)) for s in self.init_scores]
        else:
            doc["init_score"] = float(np.float32(self.init_score))
        return doc

    def bin_space(self, cls: int = 0) -> list:
        """The trees of class `cls` (tree t belongs to class t % n_classes) as the device
        predictor's bin-space nodes (NODE_DTYPE)."""
        from . import NODE_DTYPE
        out = []
        for t, tree in enumerate(self.trees):
            if t % self.n_classes != cls:
                continue
            if not isinstance(tree, DenseTree):
                raise ModelFormatError("only DenseTree models convert to bin space")
            nodes = np.zeros(len(tree.feature_id), NODE_DTYPE)
            for i in range(len(nodes)):
                f = int(tree.feature_id[i])
                nodes[i]["cover"] = tree.covers[i]
                nodes[i]["gain"] = tree.gains[i]
                if f == LEAF_FLAG:
                    nodes[i]["feature"] = -1
                    nodes[i]["value"] = tree.value[i]
                    continue
                cuts = self.cuts[f]
                b = int(np.searchsorted(cuts, tree.value[i], side="left"))
                if b >= len(cuts) or cuts[b] != tree.value[i]:
                    raise ModelFormatError(f"tree {t} node {i}: threshold is not a cut of feature {f}")
                nodes[i]["feature"] = f
                nodes[i]["bin"] = b
                nodes[i]["default_left"] = int(bool(tree.default_left[i]))
                nodes[i]["left"] = tree.left[i]
                nodes[i]["right"] = tree.right[i]
            out.append(nodes)
        return out


def from_document(doc) -> Model:
    if not isinstance(doc, dict) or doc.get("magic") != MAGIC:
        raise ModelFormatError("model: bad magic")
    if doc.get("version") != FORMAT_VERSION:
        raise ModelFormatError("model: unsupported format version")
    mappers = doc["bin_mappers"]
    trees = [DenseTree.from_record(r) if "nodes" in r else r for r in doc["trees"]]
    return Model(config=doc["config"], feature_names=[m["name"] for m in mappers],
                 cuts=[np.array(m["cuts"], np.float32) for m in mappers],
                 init_score=doc.get("init_score"), init_scores=doc.get("init_scores"), trees=trees)


def loads(data: bytes) -> Model:
    return from_document(unpack(data))


def dumps(model: Model) -> bytes:
    return pack(model.document())


def load(path) -> Model:
    """io::load_booster: read a v7 file."""
    with open(path, "rb") as f:
        return loads(f.read())


def save(model: Model, path) -> None:
    """io::save_booster: write a v7 file."""
    with open(path, "wb") as f:
        f.write(dumps(model))


def from_fit(trees: list, cuts: list, config: dict, init_score: float, feature_names: list | None = None) -> Model:
    """A model of this package's fit (bin-space trees from HipContext.trees()) in the v7
    layout: split bin b becomes threshold cuts[f][b]; node ids are kept."""
    out = []
    for nodes in trees:
        n = len(nodes)
        leaf = nodes["feature"] < 0
        feat = np.where(leaf, LEAF_FLAG, nodes["feature"]).astype(np.uint32)
        val = np.array([nodes["value"][i] if leaf[i] else cuts[nodes["feature"][i]][nodes["bin"][i]]
                        for i in range(n)], np.float32)
        depth = _depth(nodes)
        out.append(DenseTree(feat, val, np.where(leaf, 0, nodes["left"]).astype(np.uint32),
                             np.where(leaf, 0, nodes["right"]).astype(np.uint32),
                             np.where(leaf, False, nodes["default_left"] != 0),
                             np.where(leaf, 0.0, nodes["gain"]).astype(np.float32),
                             nodes["cover"].astype(np.float32), depth, int(leaf.sum())))
    names = feature_names or [f"f{i}" for i in range(len(cuts))]
    return Model(config=config, feature_names=names, cuts=[np.asarray(c, np.float32) for c in cuts],
                 init_score=float(init_score), trees=out)


def _depth(nodes) -> int:
    best = 0
    stack = [(0, 0)]
    while stack:
        i, d = stack.pop()
        if nodes["feature"][i] < 0:
            best = max(best, d)
        else:
            stack.append((int(nodes["left"][i]), d + 1))
            stack.append((int(nodes["right"][i]), d + 1))
    return best
