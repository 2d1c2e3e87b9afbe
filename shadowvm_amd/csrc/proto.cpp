// proto.cpp — proto3 wire-format codec for the plan messages of com/svm.proto:73-144
// (ColumnInfo, ObjectInfo, OperatorInfo with its map<string,string> args, PipelineInfo).
// libprotobuf is not a dependency: the few messages on the path are decoded by hand.
#include "svm_internal.h"

namespace svm {
namespace {

struct Reader {
  const uint8_t* p;
  const uint8_t* end;
  bool done() const { return p >= end; }
  uint64_t varint() {
    uint64_t v = 0;
    for (int shift = 0; shift < 64; shift += 7) {
      if (p >= end) fail(SVM_ERR_INVALID, "PipelineInfo: truncated varint");
      uint8_t b = *p++;
      v |= (uint64_t)(b & 0x7F) << shift;
      if (!(b & 0x80)) return v;
    }
    fail(SVM_ERR_INVALID, "PipelineInfo: varint too long");
  }
  Reader sub() {
    uint64_t n = varint();
    if (n > (uint64_t)(end - p)) fail(SVM_ERR_INVALID, "PipelineInfo: truncated field");
    Reader r{p, p + n};
    p += n;
    return r;
  }
  std::string str() {
    Reader r = sub();
    return std::string((const char*)r.p, (size_t)(r.end - r.p));
  }
  void skip(int wt) {
    switch (wt) {
      case 0: varint(); break;
      case 1: advance(8); break;
      case 2: sub(); break;
      case 5: advance(4); break;
      default: fail(SVM_ERR_INVALID, "PipelineInfo: unsupported wire type");
    }
  }
  // occurrences of length-delimited field f at this level (to size vectors once)
  size_t count(int f) const {
    Reader r{p, end};
    size_t n = 0;
    while (!r.done()) {
      uint64_t tag = r.varint();
      int g = (int)(tag >> 3), wt = (int)(tag & 7);
      if (g == f && wt == 2) ++n;
      r.skip(wt);
    }
    return n;
  }
  void advance(size_t n) {
    if (n > (size_t)(end - p)) fail(SVM_ERR_INVALID, "PipelineInfo: truncated fixed field");
    p += n;
  }
};

ColumnInfo read_column(Reader r) {
  ColumnInfo c;
  c.val_size = 0;
  c.type = 0;
  while (!r.done()) {
    uint64_t tag = r.varint();
    int f = (int)(tag >> 3), wt = (int)(tag & 7);
    if (f == 1 && wt == 2) c.id = r.str();
    else if (f == 2 && wt == 0) c.type = (int)r.varint();
    else if (f == 3 && wt == 0) c.isnull = r.varint() != 0;
    else if (f == 4 && wt == 2) {
      Reader pk = r.sub();
      while (!pk.done()) c.dim.push_back((int32_t)pk.varint());
    } else if (f == 4 && wt == 0) c.dim.push_back((int32_t)r.varint());
    else if (f == 5 && wt == 0) c.val_size = (int64_t)r.varint();
    else r.skip(wt);
  }
  return c;
}

ObjectInfo read_object(Reader r) {
  ObjectInfo o;
  o.column_list.reserve(r.count(3));
  while (!r.done()) {
    uint64_t tag = r.varint();
    int f = (int)(tag >> 3), wt = (int)(tag & 7);
    if (f == 1 && wt == 2) o.id = r.str();
    else if (f == 2 && wt == 2) o.host = r.str();
    else if (f == 3 && wt == 2) o.column_list.push_back(read_column(r.sub()));
    else if (f == 4 && wt == 0) o.storage = (int)r.varint();
    else if (f == 5 && wt == 2) o.path = r.str();
    else r.skip(wt);
  }
  return o;
}

OperatorInfo read_operator(Reader r) {
  OperatorInfo op;
  bool typed = false;
  while (!r.done()) {
    uint64_t tag = r.varint();
    int f = (int)(tag >> 3), wt = (int)(tag & 7);
    if (f == 1 && wt == 2) op.id = r.str();
    else if (f == 2 && wt == 0) {
      op.type = (int)r.varint();
      typed = true;
    }
    else if (f == 4 && wt == 2) {  // map<string,string> entry {1: key, 2: value}
      Reader e = r.sub();
      std::string k, v;
      while (!e.done()) {
        uint64_t t2 = e.varint();
        int f2 = (int)(t2 >> 3), w2 = (int)(t2 & 7);
        if (f2 == 1 && w2 == 2) k = e.str();
        else if (f2 == 2 && w2 == 2) v = e.str();
        else e.skip(w2);
      }
      op.args[k] = v;  // later entries win, as protobuf map parsing does
    } else if (f == 5 && wt == 2) {
      Reader o = r.sub();
      op.has_output = true;
      // decoded when the operator's type (field 2, written first by proto3 encoders)
      // makes the schema one the plan reads, or is not known yet
      if (!typed || op.type == OpSource || op.type == OpLoader || op.type == OpSink) {
        op.output = read_object(o);
        op.output_decoded = true;
        op.output_raw.clear();
      } else {
        op.output_raw.assign((const char*)o.p, (size_t)(o.end - o.p));
        op.output_decoded = false;
      }
    } else if (f == 6 && wt == 0) op.perf = (int)r.varint();
    else r.skip(wt);  // field 3 (DependencyInfo) is not used by the data plane
  }
  return op;
}

void put_varint(std::string& s, uint64_t v) {
  while (v >= 0x80) { s.push_back((char)(v | 0x80)); v >>= 7; }
  s.push_back((char)v);
}
void put_tag(std::string& s, int f, int wt) { put_varint(s, ((uint64_t)f << 3) | (uint64_t)wt); }
void put_bytes(std::string& s, int f, const std::string& b) {
  put_tag(s, f, 2);
  put_varint(s, b.size());
  s += b;
}
void put_uint(std::string& s, int f, uint64_t v) {
  if (!v) return;
  put_tag(s, f, 0);
  put_varint(s, v);
}
std::string enc_column(const ColumnInfo& c) {
  std::string s;
  if (!c.id.empty()) put_bytes(s, 1, c.id);
  put_uint(s, 2, (uint64_t)c.type);
  put_uint(s, 3, c.isnull ? 1 : 0);
  if (!c.dim.empty()) {
    std::string pk;
    for (int32_t d : c.dim) put_varint(pk, (uint64_t)(int64_t)d);
    put_bytes(s, 4, pk);
  }
  put_uint(s, 5, (uint64_t)c.val_size);
  return s;
}
std::string enc_object(const ObjectInfo& o) {
  std::string s;
  if (!o.id.empty()) put_bytes(s, 1, o.id);
  if (!o.host.empty()) put_bytes(s, 2, o.host);
  for (auto& c : o.column_list) put_bytes(s, 3, enc_column(c));
  put_uint(s, 4, (uint64_t)o.storage);
  if (!o.path.empty()) put_bytes(s, 5, o.path);
  return s;
}

}  // namespace

PipelineInfo decode_pipeline_info(const uint8_t* data, size_t len) {
  if (!data && len) fail(SVM_ERR_INVALID, "PipelineInfo: null buffer");
  Reader r{data, data + len};
  PipelineInfo p;
  p.workflow.reserve(r.count(2));
  while (!r.done()) {
    uint64_t tag = r.varint();
    int f = (int)(tag >> 3), wt = (int)(tag & 7);
    if (f == 1 && wt == 2) p.id = r.str();
    else if (f == 2 && wt == 2) p.workflow.push_back(read_operator(r.sub()));
    else if (f == 3 && wt == 2) p.pre_pipeline_id.push_back(r.str());
    else if (f == 4 && wt == 0) p.codegen = r.varint() != 0;
    else if (f == 5 && wt == 0) p.debug = r.varint() != 0;
    else if (f == 6 && wt == 0) p.isGPU = r.varint() != 0;
    else r.skip(wt);
  }
  return p;
}

std::string encode_pipeline_info(const PipelineInfo& p) {
  std::string s;
  if (!p.id.empty()) put_bytes(s, 1, p.id);
  for (auto& op : p.workflow) {
    std::string o;
    if (!op.id.empty()) put_bytes(o, 1, op.id);
    put_uint(o, 2, (uint64_t)op.type);
    for (auto& kv : op.args) {
      std::string e;
      put_bytes(e, 1, kv.first);
      put_bytes(e, 2, kv.second);
      put_bytes(o, 4, e);
    }
    if (op.has_output) put_bytes(o, 5, op.output_decoded ? enc_object(op.output) : op.output_raw);
    put_uint(o, 6, (uint64_t)op.perf);
    put_bytes(s, 2, o);
  }
  for (auto& id : p.pre_pipeline_id) put_bytes(s, 3, id);
  put_uint(s, 4, p.codegen);
  put_uint(s, 5, p.debug);
  put_uint(s, 6, p.isGPU);
  return s;
}

}  // namespace svm
