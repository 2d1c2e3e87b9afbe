"""Wire-compatible Python classes for the plan messages of com/svm.proto.

protoc is not available, so the file descriptor of ``svm.proto`` (package
``ShadowVM``) is assembled here field by field with the same names, numbers and
types as com/svm.proto:49-159, and message classes are created from it with the
protobuf runtime.  Bytes produced by these classes are what the Spark vRDD
(java_vRDD/.../vRDD.scala:57-63) and the C++ example drivers put on the wire.
"""
from __future__ import annotations

from google.protobuf import descriptor_pb2, descriptor_pool, message_factory

_F = descriptor_pb2.FieldDescriptorProto


def _field(msg, name, number, ftype, label=_F.LABEL_OPTIONAL, type_name=None):
    f = msg.field.add()
    f.name, f.number, f.type, f.label = name, number, ftype, label
    if type_name:
        f.type_name = type_name
    return f


def _enum(container, name, values):
    e = container.enum_type.add()
    e.name = name
    for n, v in values:
        ev = e.value.add()
        ev.name, ev.number = n, v
    return e


def _build_file():
    fd = descriptor_pb2.FileDescriptorProto()
    fd.name = "svm.proto"
    fd.package = "ShadowVM"
    fd.syntax = "proto3"
    _enum(fd, "StoType", [("Disk", 0), ("HostMem", 1), ("DeviceMem", 2), ("SharedMem", 3)])
    _enum(fd, "ValType", [("Byte", 0), ("Int16", 1), ("Int32", 2), ("Int64", 3), ("Float16", 4),
                          ("Float32", 5), ("Float64", 6), ("String", 7), ("Tensor", 8)])
    R = _F.LABEL_REPEATED

    m = fd.message_type.add(); m.name = "ColumnInfo"
    _field(m, "id", 1, _F.TYPE_STRING)
    _field(m, "type", 2, _F.TYPE_ENUM, type_name=".ShadowVM.ValType")
    _field(m, "isnull", 3, _F.TYPE_BOOL)
    _field(m, "dim", 4, _F.TYPE_INT32, R)
    _field(m, "val_size", 5, _F.TYPE_INT64)

    m = fd.message_type.add(); m.name = "ObjectInfo"
    _field(m, "id", 1, _F.TYPE_STRING)
    _field(m, "host", 2, _F.TYPE_STRING)
    _field(m, "column_list", 3, _F.TYPE_MESSAGE, R, ".ShadowVM.ColumnInfo")
    _field(m, "storage", 4, _F.TYPE_ENUM, type_name=".ShadowVM.StoType")
    _field(m, "path", 5, _F.TYPE_STRING)

    m = fd.message_type.add(); m.name = "DependencyInfo"
    _enum(m, "Type", [("Whole", 0), ("HashPartition", 1), ("RangePartition", 2)])
    _field(m, "pre_object", 1, _F.TYPE_MESSAGE, type_name=".ShadowVM.ObjectInfo")
    _field(m, "type", 2, _F.TYPE_ENUM, type_name=".ShadowVM.DependencyInfo.Type")
    _field(m, "shard_id", 3, _F.TYPE_INT32)

    m = fd.message_type.add(); m.name = "OperatorInfo"
    _enum(m, "Type", [("Source", 0), ("Loader", 1), ("Filter", 2), ("Map", 3), ("ReduceBy", 4),
                      ("SortBy", 5), ("JoinWith", 6), ("Sink", 7)])
    _enum(m, "Preference", [("CPU", 0), ("GPU", 1), ("NO", 2)])
    entry = m.nested_type.add()
    entry.name = "ArgsEntry"
    entry.options.map_entry = True
    _field(entry, "key", 1, _F.TYPE_STRING)
    _field(entry, "value", 2, _F.TYPE_STRING)
    _field(m, "id", 1, _F.TYPE_STRING)
    _field(m, "type", 2, _F.TYPE_ENUM, type_name=".ShadowVM.OperatorInfo.Type")
    _field(m, "input", 3, _F.TYPE_MESSAGE, R, ".ShadowVM.DependencyInfo")
    _field(m, "args", 4, _F.TYPE_MESSAGE, R, ".ShadowVM.OperatorInfo.ArgsEntry")
    _field(m, "output", 5, _F.TYPE_MESSAGE, type_name=".ShadowVM.ObjectInfo")
    _field(m, "perf", 6, _F.TYPE_ENUM, type_name=".ShadowVM.OperatorInfo.Preference")

    m = fd.message_type.add(); m.name = "PipelineInfo"
    _field(m, "id", 1, _F.TYPE_STRING)
    _field(m, "workflow", 2, _F.TYPE_MESSAGE, R, ".ShadowVM.OperatorInfo")
    _field(m, "pre_pipeline_id", 3, _F.TYPE_STRING, R)
    _field(m, "codegen", 4, _F.TYPE_BOOL)
    _field(m, "debug", 5, _F.TYPE_BOOL)
    _field(m, "isGPU", 6, _F.TYPE_BOOL)

    m = fd.message_type.add(); m.name = "Ret"
    _field(m, "success", 1, _F.TYPE_BOOL)
    _field(m, "info", 2, _F.TYPE_STRING)
    return fd


_pool = descriptor_pool.DescriptorPool()
_pool.Add(_build_file())


def _cls(name):
    return message_factory.GetMessageClass(_pool.FindMessageTypeByName("ShadowVM." + name))


ColumnInfo = _cls("ColumnInfo")
ObjectInfo = _cls("ObjectInfo")
DependencyInfo = _cls("DependencyInfo")
OperatorInfo = _cls("OperatorInfo")
PipelineInfo = _cls("PipelineInfo")
Ret = _cls("Ret")


def to_message(plan: dict):
    """Plan dict (see shadowvm_amd.pipeline) -> PipelineInfo message."""
    p = PipelineInfo()
    p.id = plan.get("id", "")
    for op in plan["workflow"]:
        o = p.workflow.add()
        o.id = op.get("id", p.id)
        o.type = op["type"]
        for k, v in op.get("args", {}).items():
            o.args[k] = v
        out = op.get("output")
        if out is not None:
            for c in out.get("column_list", []):
                col = o.output.column_list.add()
                col.type = c["type"]
                col.val_size = c["val_size"]
            o.output.storage = out.get("storage", 1)
            o.output.id = out.get("id", "")
    p.pre_pipeline_id.extend(plan.get("pre_pipeline_id", []))
    p.codegen = plan.get("codegen", True)
    p.debug = plan.get("debug", False)
    p.isGPU = plan.get("isGPU", True)
    return p


def serialize(plan: dict) -> bytes:
    return to_message(plan).SerializeToString()
