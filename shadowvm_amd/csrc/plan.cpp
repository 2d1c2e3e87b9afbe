// plan.cpp — PipelineInfo -> operator chain, and expression lowering to HIP.
#include "plan.h"

#include <cctype>
#include <climits>
#include <cstring>

namespace svm {

std::vector<std::string> str_split(const std::string& code, const std::string& sep) {
  // com/cpp/com.cu:79-93: pieces between separators; a trailing empty piece and
  // an empty input give nothing.
  std::vector<std::string> ret;
  size_t pos1 = 0, pos2 = code.find(sep);
  while (pos2 != std::string::npos) {
    ret.push_back(code.substr(pos1, pos2 - pos1));
    pos1 = pos2 + sep.size();
    pos2 = code.find(sep, pos1);
  }
  if (pos1 != code.size()) ret.push_back(code.substr(pos1));
  return ret;
}

static int stoi_prefix(const std::string& s, const std::string& ctx) {
  // std::stoi semantics: leading blanks, optional sign, digits; the rest ignored.
  size_t k = 0;
  while (k < s.size() && isspace((unsigned char)s[k])) k++;
  size_t b = k;
  if (k < s.size() && (s[k] == '+' || s[k] == '-')) k++;
  size_t d = k;
  while (k < s.size() && isdigit((unsigned char)s[k])) k++;
  if (k == d) fail(SVM_ERR_INVALID, "bad operand index in '" + ctx + "'");
  long v = strtol(s.substr(b, k - b).c_str(), nullptr, 10);
  if (v < INT_MIN || v > INT_MAX) fail(SVM_ERR_INVALID, "operand index out of range in '" + ctx + "'");
  return (int)v;
}

KeyIndex gen_key_index(const std::string& exp) {
  // codegen/codegen.cpp:231-238: "$(o.c)" with trailing text tolerated.
  size_t p1 = exp.find("$(");
  if (p1 == std::string::npos) fail(SVM_ERR_INVALID, "key '" + exp + "' has no $(obj.col)");
  size_t p2 = exp.find('.', p1);
  if (p2 == std::string::npos) fail(SVM_ERR_INVALID, "key '" + exp + "' has no '.'");
  KeyIndex k;
  k.obj = stoi_prefix(exp.substr(p1 + 2, p2 - p1 - 2), exp);
  k.col = stoi_prefix(exp.substr(p2 + 1), exp);
  return k;
}

static const std::string* arg(const OperatorInfo& op, const char* name) {
  auto it = op.args.find(name);
  return it == op.args.end() ? nullptr : &it->second;
}

Plan decode_plan(const PipelineInfo& info) {
  // core/pipeline.cu:11-80: first pass Source / Loader / Sink, second pass the
  // transform operators in workflow order.
  Plan p;
  p.id = info.id;
  for (auto& op : info.workflow) {
    switch (op.type) {
      case OpSource: {
        p.has_source = true;
        const std::string* t = arg(op, "type");
        p.source_type = (t && !t->empty() && (*t)[0] == 'f') ? "file" : "object";
        const std::string* path = arg(op, "path");
        if (path) p.source_path = *path;
        p.source_info = op.output;
        break;
      }
      case OpLoader: {
        // core/operator.cpp:13-21: the separators are the first characters of the args
        p.has_loader = true;
        const std::string* cs = arg(op, "col_spilter");
        const std::string* rs = arg(op, "row_spilter");
        if (cs) p.col_sep = *cs;
        if (rs) p.row_sep = *rs;
        p.loader_info = op.output;
        break;
      }
      case OpSink:
        p.has_sink = true;
        p.output_info = op.output;
        break;
      default: break;
    }
  }
  for (auto& op : info.workflow) {
    switch (op.type) {
      case OpMap: {
        const std::string* s = arg(op, "select");
        if (!s) fail(SVM_ERR_INVALID, "Map without 'select'");
        p.select_exp = str_split(*s, ",");
        break;
      }
      case OpFilter: {
        const std::string* s = arg(op, "predication");
        if (!s) fail(SVM_ERR_INVALID, "Filter without 'predication'");
        p.filter_exp.push_back(*s);
        break;
      }
      case OpReduceBy: {
        const std::string* m = arg(op, "metric");
        if (!m) fail(SVM_ERR_INVALID, "ReduceBy without 'metric'");
        p.metric_exp = str_split(*m, ",");
        if (const std::string* d = arg(op, "dim"))
          for (auto& e : str_split(*d, ",")) p.reduce_key.push_back(gen_key_index(e));
        p.is_reduce = true;
        break;
      }
      case OpJoinWith: {
        const std::string* l = arg(op, "lkey");
        const std::string* r = arg(op, "rkey");
        if (!l || !r) fail(SVM_ERR_INVALID, "JoinWith needs 'lkey' and 'rkey'");
        p.join_lkey.push_back(gen_key_index(*l));
        p.join_rkey.push_back(gen_key_index(*r));
        break;
      }
      case OpSortBy:  // ignored by the reference (core/pipeline.cu:69-71)
      default: break;
    }
  }
  return p;
}

// ------------------------------------------------------------------ lowering
static const char* native(int t) {
  switch (t) {
    case Byte: return "char";
    case Int16: return "short";
    case Int32: return "int";
    case Int64: return "long long";
    case String: return "char";
    default: return nullptr;
  }
}

// Identifiers an expression may use: the reference runtime's functions
// (com/cpp/cpu-runtime.hpp:37-78), C++'s alternative operator spellings (the Spark
// control plane writes "and", java_vRDD/src/test/java/SSBHelper.java:371), boolean
// literals and integer type names for casts.  Anything else would name the generated
// kernel's own variables or functions, and is refused.
static bool allowed_identifier(const std::string& id) {
  static const char* const kAllowed[] = {"StrCmp", "Hash", "StrHash", "true", "false", "and", "or", "not",
                                         "not_eq", "bitand", "bitor", "xor", "compl", "char", "short", "int",
                                         "long", "signed", "unsigned"};
  for (const char* a : kAllowed)
    if (id == a) return true;
  return false;
}

// A numeric token must be a C integer or floating literal.
static bool numeric_literal(const std::string& t) {
  size_t k = 0, n = t.size();
  if (n > 2 && t[0] == '0' && (t[1] == 'x' || t[1] == 'X')) {
    k = 2;
    while (k < n && isxdigit((unsigned char)t[k])) k++;
    if (k == 2) return false;
  } else {
    while (k < n && isdigit((unsigned char)t[k])) k++;
    bool frac = false;
    if (k < n && t[k] == '.') {
      frac = true;
      k++;
      while (k < n && isdigit((unsigned char)t[k])) k++;
    }
    if (k < n && (t[k] == 'e' || t[k] == 'E')) {
      frac = true;
      k++;
      if (k < n && (t[k] == '+' || t[k] == '-')) k++;
      const size_t d = k;
      while (k < n && isdigit((unsigned char)t[k])) k++;
      if (k == d) return false;
    }
    if (frac) return k == n || (k + 1 == n && (t[k] == 'f' || t[k] == 'F'));
  }
  while (k < n && (t[k] == 'u' || t[k] == 'U' || t[k] == 'l' || t[k] == 'L')) k++;
  return k == n;
}

Lowered lower_expr(const std::string& exp, const Schemas& schemas, bool is_reduce_key,
                   Literals* lits) {
  Lowered out;
  std::string& o = out.code;
  size_t n = exp.size(), k = 0;
  // Open parentheses: whether each belongs to a StrCmp call, and the commas seen at its
  // level.  StrCmp's length (third argument) stays a literal in the generated code, so
  // the comparison unrolls over a width known when the kernel is compiled.
  std::vector<std::pair<bool, int>> parens;
  bool after_strcmp = false;
  while (k < n) {
    char ch = exp[k];
    if (ch == '(' && !(k > 0 && exp[k - 1] == '$')) {
      parens.push_back({after_strcmp, 0});
      after_strcmp = false;
      o.push_back(ch);
      k++;
      continue;
    }
    if (ch == ')' && !parens.empty()) {
      parens.pop_back();
      o.push_back(ch);
      k++;
      continue;
    }
    if (ch == ',' && !parens.empty()) {
      parens.back().second++;
      o.push_back(ch);
      k++;
      continue;
    }
    if (!isspace((unsigned char)ch) && !isalpha((unsigned char)ch) && ch != '_') after_strcmp = false;
    if (ch == '$' && k + 1 < n && exp[k + 1] == '(') {
      size_t p2 = exp.find('.', k);
      size_t p3 = p2 == std::string::npos ? p2 : exp.find(')', p2);
      if (p3 == std::string::npos) fail(SVM_ERR_INVALID, "unterminated operand in '" + exp + "'");
      int obj = stoi_prefix(exp.substr(k + 2, p2 - k - 2), exp);
      int col = stoi_prefix(exp.substr(p2 + 1, p3 - p2 - 1), exp);
      if (obj < 0 || obj >= (int)schemas.size() || col < 0 || col >= (int)schemas[obj].size())
        fail(SVM_ERR_INVALID, "operand $(" + std::to_string(obj) + "." + std::to_string(col) +
                                  ") does not exist in '" + exp + "'");
      int t = schemas[obj][col].first;
      int64_t size = schemas[obj][col].second;
      const char* nt = native(t);
      if (!nt) fail(SVM_ERR_UNSUPPORTED, "floating point / tensor operands are not on the integer path");
      out.operands.insert({obj, col});
      std::string ptr, ix = obj == 0 ? "i" : "JI" + std::to_string(obj);
      ptr = (obj == 0 ? "FCOL" : "DCOL" + std::to_string(obj) + "_") + std::to_string(col);
      if (obj == 0) ptr = "FCOL" + std::to_string(col);
      std::string pos = "(" + ptr + "+(long long)" + ix + "*" + std::to_string(size) + ")";
      if (t == String) {
        o += is_reduce_key ? "(int)Hash(" + pos + "," + std::to_string(size) + ")"
                           : "((const char*)" + pos + ")";
      } else if (obj == 0 && t == Int32) {
        o += "(f" + std::to_string(col) + ")";
      } else {
        o += "(*(const " + std::string(nt) + "*)" + pos + ")";
      }
      k = p3 + 1;
      continue;
    }
    if (ch == '"') {
      size_t e = k + 1;
      bool escaped = false;
      while (e < n && exp[e] != '"') {
        if (exp[e] == '\\') { escaped = true; e++; }
        e++;
      }
      if (e >= n) fail(SVM_ERR_INVALID, "unterminated string literal in '" + exp + "'");
      std::string body = exp.substr(k + 1, e - k - 1);
      if (escaped || !lits) {
        o += exp.substr(k, e - k + 1);
      } else {
        size_t off = lits->blob.size();
        lits->blob += body;
        lits->blob.push_back('\0');
        while (lits->blob.size() % 8) lits->blob.push_back('\0');
        o += "PS(" + std::to_string(off) + ")";
      }
      k = e + 1;
      continue;
    }
    if (ch == '\'') {
      size_t e = k + 1;
      while (e < n && exp[e] != '\'') {
        if (exp[e] == '\\') e++;
        e++;
      }
      o += exp.substr(k, e - k + 1);
      k = e + 1;
      continue;
    }
    if (isdigit((unsigned char)ch)) {
      size_t e = k;
      while (e < n && (isalnum((unsigned char)exp[e]) || exp[e] == '_' || exp[e] == '.')) e++;
      std::string tok = exp.substr(k, e - k);
      if (!numeric_literal(tok)) fail(SVM_ERR_INVALID, "malformed number '" + tok + "' in '" + exp + "'");
      bool plain = true;
      for (char c : tok) plain = plain && isdigit((unsigned char)c);
      const bool strcmp_len = !parens.empty() && parens.back().first && parens.back().second == 2;
      if (plain && (tok.size() == 1 || tok[0] != '0') && tok.size() <= 10 && lits && !strcmp_len) {
        long long v = atoll(tok.c_str());
        if (v <= INT_MAX) {
          o += "PI(" + std::to_string(lits->ints.size()) + ")";
          lits->ints.push_back((int32_t)v);
          k = e;
          continue;
        }
      }
      o += tok;  // suffixed, octal, hex, wide or float literals keep their C meaning verbatim
      k = e;
      continue;
    }
    if (isalpha((unsigned char)ch) || ch == '_') {
      size_t e = k;
      while (e < n && (isalnum((unsigned char)exp[e]) || exp[e] == '_')) e++;
      const std::string id = exp.substr(k, e - k);
      if (!allowed_identifier(id))
        fail(SVM_ERR_INVALID, "identifier '" + id + "' is not allowed in '" + exp +
                                  "' (operands are $(obj.col); functions StrCmp, Hash, StrHash)");
      o += id;
      after_strcmp = id == "StrCmp";
      k = e;
      continue;
    }
    // Operators and punctuation of C expressions only: no statements (; { }), no
    // member access or indexing (. [ ] ->), no comments, no assignments or
    // increments, no preprocessor text.
    if (strchr(";{}[]#\\@`.", ch))
      fail(SVM_ERR_INVALID, std::string("character '") + ch + "' is not allowed in '" + exp + "'");
    if (!isspace((unsigned char)ch) && !strchr("+-*/%<>=!&|^~?:,()", ch))
      fail(SVM_ERR_INVALID, "character code " + std::to_string((unsigned char)ch) + " is not allowed in '" + exp + "'");
    const char nx = k + 1 < n ? exp[k + 1] : '\0';
    if (ch == '/' && (nx == '/' || nx == '*')) fail(SVM_ERR_INVALID, "comments are not allowed in '" + exp + "'");
    if ((ch == '+' && nx == '+') || (ch == '-' && (nx == '-' || nx == '>')))
      fail(SVM_ERR_INVALID, "increments and member access are not allowed in '" + exp + "'");
    if (ch == '=') {
      // '=' only inside ==, !=, <= and >= (not <<= or >>=, shift assignments)
      const char pv = k > 0 ? exp[k - 1] : '\0';
      const bool compare = nx == '=' || (pv && strchr("=!<>", pv));
      const bool shift_assign = k >= 2 && (pv == '<' || pv == '>') && exp[k - 2] == pv;
      if (!compare || shift_assign) fail(SVM_ERR_INVALID, "assignments are not allowed in '" + exp + "'");
    }
    o.push_back(ch);
    k++;
  }
  return out;
}

}  // namespace svm
