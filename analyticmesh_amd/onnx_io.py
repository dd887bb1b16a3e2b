"""ONNX load/save for ReLU MLPs without the ``onnx`` package.

The reference parses ``onnx.helper.printable_graph`` text with regular
expressions (backend/onnx_io.py:11-259) and saves with ``torch.onnx.export``
(backend/onnx_io.py:262-265).  Neither dependency exists here, so this module
reads and writes the protobuf wire format itself.  It keeps the reference's
graph rules:

* accepted ops: Gemm, Relu, MatMul, Add, Transpose (backend/onnx_io.py:100);
* Gemm must have alpha=1, beta=1, transB=1 (backend/onnx_io.py:105-108);
* a Transpose(perm=[1,0]) must be followed by the MatMul that consumes it,
  and is folded into it (backend/onnx_io.py:109-123);
* the hidden widths come from the Gemm weight shapes (backend/onnx_io.py:132-160);
* every Add that is not the bias of a Gemm is a skip connection whose source
  is a Relu output (identity) or a MatMul of one (transform), and whose
  destination is the next Relu or the graph output
  (backend/onnx_io.py:163-219).

Only the ONNX message fields this path needs are decoded; unknown fields are
skipped by wire type, so newer exporters' extra metadata is tolerated.
"""
import io
import os
import struct

import numpy as np
import torch

from .model import MLP

__all__ = ["load_model", "save_model", "read_onnx_graph", "write_onnx_graph"]

# ----------------------------------------------------------------------------
# protobuf wire format
# ----------------------------------------------------------------------------
_VARINT, _I64, _LEN, _I32 = 0, 1, 2, 5


def _read_varint(buf, pos):
    result = 0
    shift = 0
    while True:
        if pos >= len(buf):
            raise ValueError("truncated varint")
        b = buf[pos]
        pos += 1
        result |= (b & 0x7F) << shift
        if not b & 0x80:
            return result, pos
        shift += 7
        if shift > 70:
            raise ValueError("varint too long")


def _fields(buf):
    """Yield (field_number, wire_type, value) for one message body."""
    pos = 0
    n = len(buf)
    while pos < n:
        key, pos = _read_varint(buf, pos)
        fno, wt = key >> 3, key & 7
        if wt == _VARINT:
            val, pos = _read_varint(buf, pos)
        elif wt == _I64:
            val = buf[pos:pos + 8]
            pos += 8
        elif wt == _LEN:
            ln, pos = _read_varint(buf, pos)
            val = buf[pos:pos + ln]
            if len(val) != ln:
                raise ValueError("truncated length-delimited field")
            pos += ln
        elif wt == _I32:
            val = buf[pos:pos + 4]
            pos += 4
        else:
            raise ValueError(f"unsupported wire type {wt}")
        yield fno, wt, val


def _signed64(v):
    return v - (1 << 64) if v >= (1 << 63) else v


def _packed_varints(val, wt):
    if wt == _VARINT:
        return [_signed64(val)]
    out = []
    pos = 0
    while pos < len(val):
        v, pos = _read_varint(val, pos)
        out.append(_signed64(v))
    return out


def _packed_floats(val, wt):
    if wt == _I32:
        return list(struct.unpack("<f", val))
    return list(struct.unpack(f"<{len(val) // 4}f", val))


def _packed_doubles(val, wt):
    if wt == _I64:
        return list(struct.unpack("<d", val))
    return list(struct.unpack(f"<{len(val) // 8}d", val))


# ONNX TensorProto.DataType values (public onnx.proto enum)
_ONNX_FLOAT, _ONNX_INT32, _ONNX_INT64, _ONNX_DOUBLE = 1, 6, 7, 11
_NP_OF = {_ONNX_FLOAT: np.float32, _ONNX_INT32: np.int32,
          _ONNX_INT64: np.int64, _ONNX_DOUBLE: np.float64}
_DTYPE_NAME = {_ONNX_FLOAT: "FLOAT", _ONNX_INT32: "INT32",
               _ONNX_INT64: "INT64", _ONNX_DOUBLE: "DOUBLE"}


def _decode_tensor(buf):
    dims, dtype, name, raw = [], 0, "", None
    fdata, ddata, i32data, i64data = [], [], [], []
    for fno, wt, val in _fields(buf):
        if fno == 1:
            dims += _packed_varints(val, wt)
        elif fno == 2:
            dtype = val
        elif fno == 4:
            fdata += _packed_floats(val, wt)
        elif fno == 5:
            i32data += _packed_varints(val, wt)
        elif fno == 7:
            i64data += _packed_varints(val, wt)
        elif fno == 8:
            name = val.decode()
        elif fno == 9:
            raw = bytes(val)
        elif fno == 10:
            ddata += _packed_doubles(val, wt)
    if dtype not in _NP_OF:
        raise ValueError(f"tensor {name!r}: unsupported data type {dtype}")
    npt = _NP_OF[dtype]
    if raw is not None:
        arr = np.frombuffer(raw, dtype=np.dtype(npt).newbyteorder("<")).astype(npt)
    elif dtype == _ONNX_FLOAT:
        arr = np.asarray(fdata, dtype=np.float32)
    elif dtype == _ONNX_DOUBLE:
        arr = np.asarray(ddata, dtype=np.float64)
    elif dtype == _ONNX_INT32:
        arr = np.asarray(i32data, dtype=np.int32)
    else:
        arr = np.asarray(i64data, dtype=np.int64)
    shape = tuple(int(d) for d in dims)
    if int(np.prod(shape)) != arr.size:
        raise ValueError(f"tensor {name!r}: {arr.size} values for shape {shape}")
    return name, _DTYPE_NAME[dtype], arr.reshape(shape)


def _decode_attribute(buf):
    att = {"name": "", "f": None, "i": None, "ints": [], "floats": [], "type": 0}
    for fno, wt, val in _fields(buf):
        if fno == 1:
            att["name"] = val.decode()
        elif fno == 2:
            att["f"] = struct.unpack("<f", val)[0]
        elif fno == 3:
            att["i"] = _signed64(val)
        elif fno == 7:
            att["floats"] += _packed_floats(val, wt)
        elif fno == 8:
            att["ints"] += _packed_varints(val, wt)
        elif fno == 20:
            att["type"] = val
    if att["type"] == 1 or (att["type"] == 0 and att["f"] is not None):
        value = att["f"]
    elif att["type"] == 2 or (att["type"] == 0 and att["i"] is not None):
        value = att["i"]
    elif att["type"] == 7:
        value = list(att["ints"])
    elif att["type"] == 6:
        value = list(att["floats"])
    else:
        value = None
    return att["name"], value


def _decode_node(buf):
    node = {"inputs": [], "outputs": [], "name": "", "op": "", "attrs": {}}
    for fno, wt, val in _fields(buf):
        if fno == 1:
            node["inputs"].append(val.decode())
        elif fno == 2:
            node["outputs"].append(val.decode())
        elif fno == 3:
            node["name"] = val.decode()
        elif fno == 4:
            node["op"] = val.decode()
        elif fno == 5:
            k, v = _decode_attribute(val)
            node["attrs"][k] = v
    return node


def _decode_value_info(buf):
    name, elem, shape = "", 0, []
    for fno, wt, val in _fields(buf):
        if fno == 1:
            name = val.decode()
        elif fno == 2:  # TypeProto
            for f2, w2, v2 in _fields(val):
                if f2 == 1:  # tensor_type
                    for f3, w3, v3 in _fields(v2):
                        if f3 == 1:
                            elem = v3
                        elif f3 == 2:  # TensorShapeProto
                            for f4, w4, v4 in _fields(v3):
                                if f4 == 1:  # Dimension
                                    d = None
                                    for f5, w5, v5 in _fields(v4):
                                        if f5 == 1:
                                            d = _signed64(v5)
                                        elif f5 == 2:
                                            d = v5.decode()
                                    shape.append(d)
    return {"name": name, "dtype": _DTYPE_NAME.get(elem, str(elem)), "shape": shape}


def read_onnx_graph(data):
    """Decode ModelProto bytes into a plain dict (nodes, initializers, io)."""
    graph_buf = None
    opset = None
    for fno, wt, val in _fields(data):
        if fno == 7:
            graph_buf = val
        elif fno == 8:
            for f2, w2, v2 in _fields(val):
                if f2 == 2:
                    opset = v2
    if graph_buf is None:
        raise ValueError("ONNX model has no graph")
    graph = {"nodes": [], "inits": {}, "inputs": [], "outputs": [], "opset": opset}
    for fno, wt, val in _fields(graph_buf):
        if fno == 1:
            graph["nodes"].append(_decode_node(val))
        elif fno == 5:
            name, dt, arr = _decode_tensor(val)
            graph["inits"][name] = (dt, arr)
        elif fno == 11:
            graph["inputs"].append(_decode_value_info(val))
        elif fno == 12:
            graph["outputs"].append(_decode_value_info(val))
    # initializers may also be listed as graph inputs (older exporters)
    graph["inputs"] = [v for v in graph["inputs"] if v["name"] not in graph["inits"]]
    return graph


# ----------------------------------------------------------------------------
# graph -> MLP (rules of backend/onnx_io.py:85-247)
# ----------------------------------------------------------------------------
def _graph_to_mlp(graph):
    inits = graph["inits"]
    if len(graph["inputs"]) != 1:
        raise ValueError("expected exactly one graph input")
    if len(graph["outputs"]) != 1:
        raise ValueError("expected exactly one graph output")
    inp = graph["inputs"][0]
    if len(inp["shape"]) != 2 or inp["shape"][1] != 3:
        raise ValueError(f"input must have shape [N, 3], got {inp['shape']}")
    if inp["dtype"] != "FLOAT":
        raise ValueError("input must be FLOAT")
    for name, (dt, _) in inits.items():
        if dt != "FLOAT":
            raise ValueError(f"parameter {name!r} must be FLOAT, got {dt}")
    ret = graph["outputs"][0]["name"]

    flows = []
    for nd in graph["nodes"]:
        if nd["op"] not in ("Gemm", "Relu", "MatMul", "Add", "Transpose"):
            raise ValueError(f"unrecognized op: {nd['op']}")
        flows.append(dict(out=nd["outputs"][0], fun_name=nd["op"],
                          fun_args=tuple(nd["inputs"]), fun_configs=nd["attrs"]))

    # check_graph (backend/onnx_io.py:85-125)
    do_transpose = False
    expect_next = None
    for f in flows:
        if expect_next:
            if f["fun_name"] != expect_next:
                raise ValueError("Transpose must be followed by MatMul")
            expect_next = None
        cfg = f["fun_configs"]
        if f["fun_name"] == "Gemm":
            if cfg.get("alpha", 1.0) != 1 or cfg.get("beta", 1.0) != 1 or cfg.get("transB", 0) != 1:
                raise ValueError("Gemm must have alpha=1, beta=1, transB=1")
            if cfg.get("transA", 0) != 0:
                raise ValueError("Gemm must have transA=0")
        elif f["fun_name"] == "Transpose":
            if list(cfg.get("perm", [])) != [1, 0]:
                raise ValueError("Transpose must have perm=[1, 0]")
            do_transpose = True
            expect_next = "MatMul"
    if do_transpose:
        for idx, f in enumerate(flows):
            if f is not None and f["fun_name"] == "MatMul":
                pre = flows[idx - 1]
                if pre is None or pre["fun_name"] != "Transpose":
                    raise ValueError("MatMul must follow a Transpose in this graph")
                f["fun_args"] = (f["fun_args"][0], pre["fun_args"][0])
                flows[idx - 1] = None
        flows = [f for f in flows if f is not None]

    def find_out(name):
        for f in flows:
            if f["out"] == name:
                return f
        return None

    def find_consumer(name):
        for f in flows:
            if name in f["fun_args"]:
                return f
        return None

    gemms = [f for f in flows if f["fun_name"] == "Gemm"]
    relus = [f for f in flows if f["fun_name"] == "Relu"]
    if len(relus) != len(gemms) - 1 or len(gemms) < 2:
        raise ValueError("expected one Relu between consecutive Gemm layers")
    for g in gemms:
        if len(g["fun_args"]) != 3:
            raise ValueError("Gemm needs a bias")
        for a in g["fun_args"][1:]:
            if a not in inits:
                raise ValueError(f"Gemm parameter {a!r} is not an initializer")
    nodes = [3] + [int(inits[g["fun_args"][1]][1].shape[0]) for g in gemms[:-1]] + [1]
    if inits[gemms[-1]["fun_args"][1]][1].shape[0] != 1:
        raise ValueError("the last Gemm must have one output")

    aux = [inp["name"]] + [r["out"] for r in relus] + [ret]
    shortcuts = []
    for f in flows:
        if f["fun_name"] != "Add":
            continue
        # destination: follow consumers until a Relu or the graph output
        dst = f["out"]
        f2 = None
        while dst != ret and (f2 is None or f2["fun_name"] != "Relu"):
            f2 = find_consumer(dst)
            if f2 is None:
                raise ValueError(f"dangling tensor {dst!r}")
            dst = f2["out"]
        dst_index = aux.index(dst)
        src_index, tm_name = None, None
        for arg in f["fun_args"]:
            f2 = find_out(arg)
            if f2 is None:
                if arg in aux:
                    src_index, tm_name = aux.index(arg), "identity_matrix"
                    break
                continue
            if f2["fun_name"] in ("Gemm", "Add"):
                continue
            if f2["fun_name"] == "Relu":
                src_index, tm_name = aux.index(arg), "identity_matrix"
                break
            if f2["fun_name"] == "MatMul":
                src_index, tm_name = aux.index(f2["fun_args"][0]), f2["fun_args"][1]
                break
        if src_index is None:
            raise ValueError("cannot determine the source of an Add")
        shortcuts.append(dict(src_index=src_index, dst_index=dst_index, tm_name=tm_name))

    arc_table = [[0] for _ in range(len(nodes) - 2)]
    arc_tm_shape, tm_names = [], []
    for sc in shortcuts:
        if sc["tm_name"] not in tm_names:
            tm_names.append(sc["tm_name"])
            if sc["tm_name"] == "identity_matrix":
                arc_tm_shape.append([0, 0])
            else:
                shp = list(inits[sc["tm_name"]][1].shape)
                arc_tm_shape.append(shp if do_transpose else [shp[1], shp[0]])
        if sc["dst_index"] - 2 < 0:
            raise ValueError("skip connection into the first hidden layer is not supported")
        # row r holds the skips added after linear layer r+1; the last row is
        # the output layer (backend/model.py:111-115)
        row = arc_table[sc["dst_index"] - 2]
        row[0] += 1
        row.extend([sc["src_index"], tm_names.index(sc["tm_name"])])

    model = MLP(nodes=nodes, arc_table=arc_table, arc_tm_shape=arc_tm_shape,
                initialization=None, enable_print=False)
    with torch.no_grad():
        for i, g in enumerate(gemms):
            model.linears[i].weight.copy_(torch.from_numpy(inits[g["fun_args"][1]][1].copy()))
            model.linears[i].bias.copy_(torch.from_numpy(inits[g["fun_args"][2]][1].copy()))
        for i, name in enumerate(tm_names):
            if name != "identity_matrix":
                w = inits[name][1].copy()
                model.tms[i].weight.copy_(torch.from_numpy(w if do_transpose else w.T.copy()))
    return model


def load_model(model_file_str):
    """Load an ONNX ReLU MLP as :class:`MLP` (backend/onnx_io.py:268-285).

    ``model_file_str`` is a path or the file's bytes.
    """
    if isinstance(model_file_str, str) and os.path.exists(model_file_str):
        with open(model_file_str, "rb") as fh:
            data = fh.read()
    elif isinstance(model_file_str, (bytes, bytearray)):
        data = bytes(model_file_str)
    else:
        raise Exception(f"Error: unknown model_file_str = {model_file_str}")
    return _graph_to_mlp(read_onnx_graph(data))


# ----------------------------------------------------------------------------
# writer
# ----------------------------------------------------------------------------
def _varint(v):
    if v < 0:
        v += 1 << 64
    out = bytearray()
    while True:
        b = v & 0x7F
        v >>= 7
        if v:
            out.append(b | 0x80)
        else:
            out.append(b)
            return bytes(out)


def _key(fno, wt):
    return _varint((fno << 3) | wt)


def _f_varint(fno, v):
    return _key(fno, _VARINT) + _varint(int(v))


def _f_bytes(fno, b):
    if isinstance(b, str):
        b = b.encode()
    return _key(fno, _LEN) + _varint(len(b)) + b


def _f_float(fno, x):
    return _key(fno, _I32) + struct.pack("<f", x)


def _tensor_proto(name, arr):
    arr = np.ascontiguousarray(np.asarray(arr, dtype=np.float32))
    body = b"".join(_f_varint(1, d) for d in arr.shape)
    body += _f_varint(2, _ONNX_FLOAT) + _f_bytes(8, name)
    body += _f_bytes(9, arr.astype("<f4").tobytes())
    return body


def _value_info(name, shape):
    dims = b"".join(_f_bytes(1, _f_varint(1, d)) for d in shape)
    tensor_type = _f_varint(1, _ONNX_FLOAT) + _f_bytes(2, dims)
    return _f_bytes(1, name) + _f_bytes(2, _f_bytes(1, tensor_type))


def _attr_int(name, v):
    return _f_bytes(1, name) + _f_varint(3, v) + _f_varint(20, 2)


def _attr_float(name, v):
    return _f_bytes(1, name) + _f_float(2, v) + _f_varint(20, 1)


def _attr_ints(name, vs):
    return _f_bytes(1, name) + b"".join(_f_varint(8, v) for v in vs) + _f_varint(20, 7)


def _node(op, inputs, outputs, name, attrs=()):
    body = b"".join(_f_bytes(1, i) for i in inputs)
    body += b"".join(_f_bytes(2, o) for o in outputs)
    body += _f_bytes(3, name) + _f_bytes(4, op)
    body += b"".join(_f_bytes(5, a) for a in attrs)
    return body


def write_onnx_graph(nodes, inits, input_name, output_name, graph_name="analytic_mlp"):
    """Serialize a node list + initializer dict as ModelProto bytes (opset 9)."""
    g = b"".join(_f_bytes(1, _node(*n)) for n in nodes)
    g += _f_bytes(2, graph_name)
    g += b"".join(_f_bytes(5, _tensor_proto(k, v)) for k, v in inits.items())
    g += _f_bytes(11, _value_info(input_name, [1, 3]))
    g += _f_bytes(12, _value_info(output_name, [1, 1]))
    m = _f_varint(1, 4) + _f_bytes(2, "analyticmesh_amd") + _f_bytes(3, "1")
    m += _f_bytes(7, g) + _f_bytes(8, _f_bytes(1, "") + _f_varint(2, 9))
    return m


def _trace_to_nodes(module):
    """Lower a torch module made of Linear/ReLU/add into ONNX nodes via torch.fx."""
    import operator
    import torch.fx

    if isinstance(module, MLP):
        gm = torch.fx.symbolic_trace(_MLPForExport(module))
    else:
        gm = torch.fx.symbolic_trace(module)
    mods = dict(gm.named_modules())
    nodes, inits = [], {}
    names = {}
    input_name, output_name = "input.1", None
    counter = [0]

    def fresh():
        counter[0] += 1
        return str(counter[0])

    for n in gm.graph.nodes:
        if n.op == "placeholder":
            names[n] = input_name
        elif n.op == "call_module" and isinstance(mods[n.target], torch.nn.Linear):
            lin = mods[n.target]
            x = names[n.args[0]]
            wname, out = f"{n.target}.weight", fresh()
            inits[wname] = lin.weight.detach().cpu().float().numpy()
            if lin.bias is not None:
                bname = f"{n.target}.bias"
                inits[bname] = lin.bias.detach().cpu().float().numpy()
                nodes.append(("Gemm", [x, wname, bname], [out], f"Gemm_{out}",
                              [_attr_float("alpha", 1.0), _attr_float("beta", 1.0), _attr_int("transB", 1)]))
            else:
                tname = fresh()
                nodes.append(("Transpose", [wname], [tname], f"Transpose_{tname}", [_attr_ints("perm", [1, 0])]))
                nodes.append(("MatMul", [x, tname], [out], f"MatMul_{out}"))
            names[n] = out
        elif (n.op == "call_module" and isinstance(mods[n.target], torch.nn.ReLU)) or \
                (n.op == "call_function" and n.target in (torch.relu, torch.nn.functional.relu)) or \
                (n.op == "call_method" and n.target == "relu"):
            out = fresh()
            nodes.append(("Relu", [names[n.args[0]]], [out], f"Relu_{out}"))
            names[n] = out
        elif n.op == "call_module" and isinstance(mods[n.target], torch.nn.Identity):
            names[n] = names[n.args[0]]
        elif n.op == "call_function" and n.target in (operator.add, torch.add):
            out = fresh()
            nodes.append(("Add", [names[n.args[0]], names[n.args[1]]], [out], f"Add_{out}"))
            names[n] = out
        elif n.op == "output":
            output_name = names[n.args[0]]
        else:
            raise ValueError(f"save_model: unsupported operation {n.op} {n.target}")
    return nodes, inits, input_name, output_name


class _MLPForExport(torch.nn.Module):
    """A trace-friendly view of :class:`MLP` (static skip wiring, no state lists)."""

    def __init__(self, mlp):
        super().__init__()
        self.linears = mlp.linears
        self.tms = mlp.tms
        self.arc_table = [list(r) for r in mlp.arc_table]
        self.n = mlp.num_of_linears

    def forward(self, x):
        srcs = {}
        needed = {r[1 + 2 * j] for r in self.arc_table for j in range(r[0])}
        for i in range(self.n):
            if i in needed:
                srcs[i] = x
            x = self.linears[i](x)
            if i >= 1 and i - 1 < len(self.arc_table):
                row = self.arc_table[i - 1]
                for j in range(row[0]):
                    x = x + self.tms[row[2 + 2 * j]](srcs[row[1 + 2 * j]])
            if i != self.n - 1:
                x = torch.relu(x)
        return x


def save_model(model, model_path):
    """Save a ReLU MLP as ONNX (backend/onnx_io.py:262-265).

    Like the reference, the module is moved to the CPU in float32 first.
    """
    model.cpu().float()
    nodes, inits, inp, out = _trace_to_nodes(model)
    data = write_onnx_graph(nodes, inits, inp, out)
    if isinstance(model_path, (str, os.PathLike)):
        with open(model_path, "wb") as fh:
            fh.write(data)
    else:
        model_path.write(data)
    return data
