// loader.cpp — the Loader operator on the device (see loader.h).
//
// The reference maps each file of the Source path as one bulk (a directory gives
// one bulk per file, object/object.cpp:122-169) and parses it on the host, one
// task per bulk, into a sink bulk (core/operator.cpp:24-48).  Here every file is
// mapped, its rows counted on the host (memchr), its bytes copied to HBM and
// parsed there; the files' rows are concatenated, in file-name order, into one
// resident table cached under the pipeline id.
#include "loader.h"

#include <dirent.h>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

#include "kernels.h"

namespace svm {

namespace {

struct Mapped {
  std::string path;
  int fd = -1;
  const char* data = nullptr;
  uint64_t size = 0;
  uint64_t rows = 0;
  ~Mapped() {
    if (data) munmap((void*)data, size);
    if (fd >= 0) close(fd);
  }
};

std::vector<std::string> source_files(const std::string& path) {
  struct stat st;
  if (path.empty() || stat(path.c_str(), &st) != 0) fail(SVM_ERR_NOT_FOUND, "Loader source '" + path + "' not found");
  std::vector<std::string> files;
  if (S_ISDIR(st.st_mode)) {
    DIR* d = opendir(path.c_str());
    if (!d) fail(SVM_ERR_NOT_FOUND, "cannot open directory '" + path + "'");
    while (struct dirent* e = readdir(d)) {
      std::string name = e->d_name;
      if (name.empty() || name[0] == '.') continue;  // as object/object.cpp:133
      std::string full = path + (path.back() == '/' ? "" : "/") + name;
      struct stat fs;
      if (stat(full.c_str(), &fs) == 0 && S_ISREG(fs.st_mode)) files.push_back(full);
    }
    closedir(d);
    std::sort(files.begin(), files.end());  // a stable order (readdir's is the file system's)
  } else {
    files.push_back(path);
  }
  return files;
}

// Occurrences of `b` in [p, p+n), counted by up to 16 threads over equal slices.
uint64_t count_byte(const char* p, uint64_t n, char b) {
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const unsigned nt = n < (8u << 20) ? 1u : hw;
  std::vector<uint64_t> part(nt, 0);
  auto work = [&](unsigned t) {
    const char* q = p + n * t / nt;
    const char* e = p + n * (t + 1) / nt;
    uint64_t c = 0;
    while ((q = (const char*)memchr(q, b, (size_t)(e - q))) != nullptr) {
      ++c;
      ++q;
    }
    part[t] = c;
  };
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; ++t) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
  uint64_t total = 0;
  for (uint64_t c : part) total += c;
  return total;
}

}  // namespace

std::unique_ptr<svm_result> run_load(const PipelineInfo& info, const Plan& plan, const RunOptions& opt) {
  Engine& E = engine();
  const auto t0 = std::chrono::steady_clock::now();
  if (plan.source_type != "file") fail(SVM_ERR_INVALID, "a Loader pipeline reads a Source of type file");
  if (plan.col_sep.empty() || plan.row_sep.empty())
    fail(SVM_ERR_INVALID, "Loader needs col_spilter and row_spilter (core/operator.cpp:16-17)");
  const char col_sep = plan.col_sep[0], row_sep = plan.row_sep[0];
  const ObjectInfo& schema = !plan.output_info.column_list.empty() ? plan.output_info
                             : !plan.loader_info.column_list.empty() ? plan.loader_info
                                                                     : plan.source_info;
  const int ncols = (int)schema.column_list.size();
  if (ncols == 0 || ncols > ld::kMaxLoadCols)
    fail(SVM_ERR_INVALID, "Loader schema must have 1.." + std::to_string(ld::kMaxLoadCols) + " columns");
  std::vector<std::pair<int, int64_t>> oschema;
  for (auto& c : schema.column_list) {
    if (c.type != Byte && c.type != Int16 && c.type != Int32 && c.type != Int64 && c.type != String)
      fail(SVM_ERR_UNSUPPORTED, "Loader column type " + std::to_string(c.type) + " is not on the integer path");
    int64_t vs = c.val_size > 0 ? c.val_size : (c.type == Int64 ? 8 : c.type == Int16 ? 2 : c.type == Byte ? 1 : 4);
    if (c.type == Int32 && vs != 4) fail(SVM_ERR_INVALID, "Int32 column with val_size != 4");
    oschema.push_back({c.type, vs});
  }
  // map the files and count their rows
  std::vector<std::unique_ptr<Mapped>> files;
  uint64_t total = 0;
  for (auto& f : source_files(plan.source_path)) {
    auto m = std::make_unique<Mapped>();
    m->path = f;
    m->fd = open(f.c_str(), O_RDONLY);
    if (m->fd < 0) fail(SVM_ERR_NOT_FOUND, "cannot open '" + f + "'");
    struct stat st;
    fstat(m->fd, &st);
    m->size = (uint64_t)st.st_size;
    if (m->size) {
      void* p = mmap(nullptr, m->size, PROT_READ, MAP_PRIVATE, m->fd, 0);
      if (p == MAP_FAILED) fail(SVM_ERR_NO_MEMORY, "cannot map '" + f + "'");
      m->data = (const char*)p;
      m->rows = count_byte(m->data, m->size, row_sep);
    }
    total += m->rows;
    files.push_back(std::move(m));
  }
  const double host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  hipStream_t s = opt.stream ? opt.stream : E.stream;
  TablePtr out = E.store.create(info.id, oschema, total);  // zeroed: fields a row lacks read 0
  out->max_tuple_count = E.cfg.tuple_per_bulk;
  // Loader::Host writes every row into the one sink bulk (core/operator.cpp:24-48)
  out->bulk0 = total;
  ld::ParseArgs a;
  memset(&a, 0, sizeof(a));
  a.col_sep = col_sep;
  a.ncols = ncols;
  for (int c = 0; c < ncols; ++c) {
    a.type[c] = oschema[c].first;
    a.size[c] = (int)oschema[c].second;
    a.col[c] = out->cols[c].dev;
  }
  uint64_t row0 = 0;
  float copy_ms = 0, parse_ms = 0;
  hipEvent_t e0, e1, e2;
  SVM_HIP(hipEventCreate(&e0));
  SVM_HIP(hipEventCreate(&e1));
  SVM_HIP(hipEventCreate(&e2));
  for (auto& m : files) {
    if (!m->size) continue;
    DevBuf buf(m->size + 64);
    SVM_HIP(hipEventRecord(e0, s));
    SVM_HIP(hipMemcpyAsync(buf.p, m->data, m->size, hipMemcpyHostToDevice, s));
    SVM_HIP(hipMemsetAsync((char*)buf.p + m->size, 0, 64, s));
    SVM_HIP(hipEventRecord(e1, s));
    const uint64_t nchunks = (m->size + ld::kChunk - 1) / ld::kChunk;
    DevBuf counts(nchunks * 4 + 4), offs(nchunks * 8 + 16), ends(m->rows * 8 + 8);
    ld::count_separators(buf.as<char>(), m->size, row_sep, counts.as<unsigned>(), nchunks, s);
    k::exclusive_scan_u32(counts.as<uint32_t>(), nchunks, offs.as<unsigned long long>(),
                          offs.as<unsigned long long>() + nchunks, s);
    ld::separator_positions(buf.as<char>(), m->size, row_sep, offs.as<unsigned long long>(),
                            ends.as<unsigned long long>(), nchunks, s);
    a.buf = buf.as<char>();
    a.n = m->size;
    a.ends = ends.as<unsigned long long>();
    a.rows = m->rows;
    a.row0 = row0;
    ld::parse_rows(a, s);
    SVM_HIP(hipEventRecord(e2, s));
    SVM_HIP(hipStreamSynchronize(s));  // before the mapping and the buffers go
    float c = 0, p = 0;
    SVM_HIP(hipEventElapsedTime(&c, e0, e1));
    SVM_HIP(hipEventElapsedTime(&p, e1, e2));
    copy_ms += c;
    parse_ms += p;
    row0 += m->rows;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  auto res = std::make_unique<svm_result>();
  res->kind = 0;
  res->rows = (int64_t)total;
  res->types = oschema;
  if (opt.want_host) {
    for (int c = 0; c < ncols; ++c) {
      std::vector<uint8_t> h(total * (uint64_t)oschema[c].second);
      if (!h.empty()) SVM_HIP(hipMemcpy(h.data(), out->cols[c].dev, h.size(), hipMemcpyDeviceToHost));
      res->cols.push_back(std::move(h));
    }
  }
  E.store.put(info.id, out);
  // timing: build = host mapping + row count, finish = host-to-HBM copies, scan = GPU parse
  res->build_ms = host_ms;
  res->finish_ms = copy_ms;
  res->scan_ms = parse_ms;
  if (E.cfg.verbose)
    fprintf(stderr, "[svm] %s: loaded %llu rows from %zu file(s): host %.3f ms, copy %.3f ms, parse %.3f ms\n",
            info.id.c_str(), (unsigned long long)total, files.size(), host_ms, copy_ms, parse_ms);
  return res;
}

}  // namespace svm
