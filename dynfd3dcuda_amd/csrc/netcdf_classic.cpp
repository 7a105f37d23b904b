// netCDF classic format reader/writer (see netcdf_classic.h).  Big-endian
// header: magic "CDF"+version, numrecs, dim_list, gatt_list, var_list; data
// of fixed-size variables at their 'begin' offsets, record variables
// interleaved per record.
#include "netcdf_classic.h"
#include <cstring>

namespace cgfdm {

namespace {

const uint32_t TAG_DIM = 0x0A, TAG_VAR = 0x0B, TAG_ATT = 0x0C;

int type_size(int t) {
  switch (t) {
    case NC_BYTE_: case NC_CHAR_: return 1;
    case NC_SHORT_: return 2;
    case NC_INT_: case NC_FLOAT_: return 4;
    case NC_DOUBLE_: return 8;
  }
  return 0;
}
long pad4(long n) { return (n + 3) & ~3L; }

struct Reader {
  FILE* f;
  bool ok = true;
  uint32_t u32() {
    unsigned char b[4];
    if (fread(b, 1, 4, f) != 4) { ok = false; return 0; }
    return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3];
  }
  uint64_t u64() { uint64_t hi = u32(); return (hi << 32) | u32(); }
  std::string name() {
    uint32_t n = u32();
    if (!ok || n > 4096) { ok = false; return ""; }
    std::string s(n, '\0');
    if (n && fread(&s[0], 1, n, f) != n) ok = false;
    long pad = pad4(n) - n;
    if (pad) fseek(f, pad, SEEK_CUR);
    return s;
  }
  void skip_atts() {
    uint32_t tag = u32(), n = u32();
    if (tag == 0 && n == 0) return;
    if (tag != TAG_ATT) { ok = false; return; }
    for (uint32_t a = 0; a < n && ok; a++) {
      name();
      int t = (int)u32();
      uint32_t ne = u32();
      fseek(f, pad4((long)ne * type_size(t)), SEEK_CUR);
    }
  }
};

struct Writer {
  std::vector<unsigned char> b;
  void u32(uint32_t v) { for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s)); }
  void u64(uint64_t v) { u32((uint32_t)(v >> 32)); u32((uint32_t)v); }
  void name(const std::string& s) {
    u32((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    while (b.size() % 4) b.push_back(0);
  }
};

void swap_copy(const void* src, unsigned char* dst, long n, int ts) {
  const unsigned char* s = (const unsigned char*)src;
  for (long q = 0; q < n; q++)
    for (int t = 0; t < ts; t++) dst[q * ts + t] = s[q * ts + (ts - 1 - t)];
}

}  // namespace

long NcFile::Var::count() const { return 0; }

NcFile::~NcFile() { close(); }

const NcFile::Var* NcFile::find(const std::string& name) const {
  for (auto& v : vars_) if (v.name == name) return &v;
  return nullptr;
}

bool NcFile::open_read(const std::string& path, std::string& err) {
  close();
  fp_ = fopen(path.c_str(), "rb");
  if (!fp_) { err = "nc error: cannot open " + path; return false; }
  writing_ = false;
  return read_header(err);
}

bool NcFile::read_header(std::string& err) {
  Reader r{fp_};
  char magic[4];
  if (fread(magic, 1, 4, fp_) != 4 || std::memcmp(magic, "CDF", 3) != 0 || (magic[3] != 1 && magic[3] != 2)) {
    err = "nc error: not a netCDF classic file";
    return false;
  }
  version_ = magic[3];
  numrecs_ = (long)r.u32();
  uint32_t tag = r.u32(), n = r.u32();
  dims_.clear();
  if (tag == TAG_DIM)
    for (uint32_t q = 0; q < n; q++) { Dim d; d.name = r.name(); d.len = (long)r.u32(); dims_.push_back(d); }
  r.skip_atts();
  tag = r.u32(); n = r.u32();
  vars_.clear();
  recsize_ = 0;
  if (tag == TAG_VAR)
    for (uint32_t q = 0; q < n && r.ok; q++) {
      Var v;
      v.name = r.name();
      uint32_t nd = r.u32();
      for (uint32_t a = 0; a < nd; a++) v.dims.push_back((int)r.u32());
      r.skip_atts();
      v.type = (int)r.u32();
      v.vsize = (long)r.u32();
      v.begin = version_ == 1 ? (long long)r.u32() : (long long)r.u64();
      v.is_record = !v.dims.empty() && v.dims[0] < (int)dims_.size() && dims_[v.dims[0]].len == 0;
      if (v.is_record) recsize_ += v.vsize;
      vars_.push_back(v);
    }
  if (!r.ok) { err = "nc error: truncated header"; return false; }
  int nrecvars = 0;
  for (auto& v : vars_) nrecvars += v.is_record;
  if (nrecvars == 1)  // a single record variable is stored unpadded
    for (auto& v : vars_)
      if (v.is_record) {
        long e = type_size(v.type);
        for (size_t a = 1; a < v.dims.size(); a++) e *= dims_[v.dims[a]].len;
        recsize_ = e;
      }
  return true;
}

bool NcFile::read_raw(const Var& v, std::vector<unsigned char>& raw, std::string& err) {
  long ts = type_size(v.type);
  long n = 1;
  for (size_t a = v.is_record ? 1 : 0; a < v.dims.size(); a++) n *= dims_[v.dims[a]].len;
  long recs = v.is_record ? numrecs_ : 1;
  raw.resize((size_t)(n * ts * recs));
  for (long rr = 0; rr < recs; rr++) {
    long long off = v.begin + (v.is_record ? rr * recsize_ : 0);
    if (fseeko(fp_, off, SEEK_SET) != 0 || fread(raw.data() + rr * n * ts, 1, (size_t)(n * ts), fp_) != (size_t)(n * ts)) {
      err = "nc error: short read of " + v.name;
      return false;
    }
  }
  return true;
}

template <class T>
static bool convert(const std::vector<unsigned char>& raw, int type, std::vector<T>& out) {
  int ts = type_size(type);
  long n = (long)raw.size() / ts;
  out.resize(n);
  for (long q = 0; q < n; q++) {
    const unsigned char* p = &raw[q * ts];
    switch (type) {
      case NC_BYTE_: out[q] = (T)(signed char)p[0]; break;
      case NC_SHORT_: out[q] = (T)(int16_t)((p[0] << 8) | p[1]); break;
      case NC_INT_: { int32_t v = (int32_t)(((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]); out[q] = (T)v; break; }
      case NC_FLOAT_: { uint32_t u = ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; float f; std::memcpy(&f, &u, 4); out[q] = (T)f; break; }
      case NC_DOUBLE_: { uint64_t u = 0; for (int t = 0; t < 8; t++) u = (u << 8) | p[t]; double f; std::memcpy(&f, &u, 8); out[q] = (T)f; break; }
      default: return false;
    }
  }
  return true;
}

bool NcFile::read_float(const std::string& var, std::vector<float>& out, std::string& err) {
  const Var* v = find(var);
  if (!v) { err = "nc error: variable " + var + " not found"; return false; }
  std::vector<unsigned char> raw;
  if (!read_raw(*v, raw, err)) return false;
  if (!convert(raw, v->type, out)) { err = "nc error: unsupported type of " + var; return false; }
  return true;
}
bool NcFile::read_double(const std::string& var, std::vector<double>& out, std::string& err) {
  const Var* v = find(var);
  if (!v) { err = "nc error: variable " + var + " not found"; return false; }
  std::vector<unsigned char> raw;
  if (!read_raw(*v, raw, err)) return false;
  if (!convert(raw, v->type, out)) { err = "nc error: unsupported type of " + var; return false; }
  return true;
}
bool NcFile::read_int(const std::string& var, std::vector<int>& out, std::string& err) {
  const Var* v = find(var);
  if (!v) { err = "nc error: variable " + var + " not found"; return false; }
  std::vector<unsigned char> raw;
  if (!read_raw(*v, raw, err)) return false;
  if (!convert(raw, v->type, out)) { err = "nc error: unsupported type of " + var; return false; }
  return true;
}

// ---------------------------------------------------------------------------
bool NcFile::create(const std::string& path, std::string& err) {
  close();
  fp_ = fopen(path.c_str(), "w+b");
  if (!fp_) { err = "nc error: cannot create " + path; return false; }
  writing_ = true;
  dims_.clear();
  vars_.clear();
  numrecs_ = 0;
  version_ = 1;
  return true;
}

int NcFile::def_dim(const std::string& name, long len) {
  dims_.push_back({name, len});
  return (int)dims_.size() - 1;
}

int NcFile::def_var(const std::string& name, int type, const std::vector<int>& dims) {
  Var v;
  v.name = name;
  v.type = type;
  v.dims = dims;
  v.is_record = !dims.empty() && dims_[dims[0]].len == 0;
  long n = type_size(type);
  for (size_t a = v.is_record ? 1 : 0; a < dims.size(); a++) n *= dims_[dims[a]].len;
  v.vsize = pad4(n);
  vars_.push_back(v);
  return (int)vars_.size() - 1;
}

bool NcFile::end_def(std::string& err) {
  // CDF-2 (64-bit offsets) when a fixed variable would start beyond 2 GiB
  long long fixed_total = 0;
  for (auto& v : vars_) if (!v.is_record) fixed_total += v.vsize;
  version_ = fixed_total > 0x7fffffffLL ? 2 : 1;
  auto header = [&](long long hdr_size) {
    Writer w;
    w.b.insert(w.b.end(), {'C', 'D', 'F', (unsigned char)version_});
    w.u32((uint32_t)numrecs_);
    if (dims_.empty()) { w.u32(0); w.u32(0); }
    else { w.u32(TAG_DIM); w.u32((uint32_t)dims_.size()); for (auto& d : dims_) { w.name(d.name); w.u32((uint32_t)d.len); } }
    w.u32(0); w.u32(0);  // no global attributes
    if (vars_.empty()) { w.u32(0); w.u32(0); }
    else {
      w.u32(TAG_VAR); w.u32((uint32_t)vars_.size());
      for (auto& v : vars_) {
        w.name(v.name);
        w.u32((uint32_t)v.dims.size());
        for (int di : v.dims) w.u32((uint32_t)di);
        w.u32(0); w.u32(0);  // no attributes
        w.u32((uint32_t)v.type);
        w.u32((uint32_t)std::min<long long>(v.vsize, 0xffffffffLL));
        if (version_ == 1) w.u32((uint32_t)v.begin); else w.u64((uint64_t)v.begin);
      }
    }
    (void)hdr_size;
    return w.b;
  };
  long long hsize = (long long)header(0).size();
  long long off = hsize;
  for (auto& v : vars_) if (!v.is_record) { v.begin = off; off += v.vsize; }
  int nrec = 0;
  for (auto& v : vars_) nrec += v.is_record;
  recsize_ = 0;
  for (auto& v : vars_)
    if (v.is_record) {
      v.begin = off + recsize_;
      long raw = type_size(v.type);
      for (size_t a = 1; a < v.dims.size(); a++) raw *= dims_[v.dims[a]].len;
      recsize_ += nrec == 1 ? raw : v.vsize;
    }
  std::vector<unsigned char> h = header(hsize);
  if (fseeko(fp_, 0, SEEK_SET) != 0 || fwrite(h.data(), 1, h.size(), fp_) != h.size()) { err = "nc error: header write"; return false; }
  // zero-fill fixed variables so the file is complete even if one is never put
  std::vector<unsigned char> zeros(1 << 16, 0);
  for (auto& v : vars_)
    if (!v.is_record) {
      fseeko(fp_, v.begin, SEEK_SET);
      long long left = v.vsize;
      while (left > 0) { size_t c = (size_t)std::min<long long>(left, (long long)zeros.size()); fwrite(zeros.data(), 1, c, fp_); left -= c; }
    }
  return true;
}

bool NcFile::put_var(int varid, const void* data) {
  if (!fp_ || varid < 0 || varid >= (int)vars_.size()) return false;
  const Var& v = vars_[varid];
  long ts = type_size(v.type);
  long n = 1;
  for (int di : v.dims) n *= dims_[di].len;
  std::vector<unsigned char> buf((size_t)(n * ts));
  swap_copy(data, buf.data(), n, (int)ts);
  return fseeko(fp_, v.begin, SEEK_SET) == 0 && fwrite(buf.data(), 1, buf.size(), fp_) == buf.size();
}

bool NcFile::put_record(int varid, long rec, const void* data) {
  if (!fp_ || varid < 0 || varid >= (int)vars_.size()) return false;
  const Var& v = vars_[varid];
  long ts = type_size(v.type);
  long n = 1;
  for (size_t a = 1; a < v.dims.size(); a++) n *= dims_[v.dims[a]].len;
  std::vector<unsigned char> buf((size_t)pad4(n * ts), 0);
  swap_copy(data, buf.data(), n, (int)ts);
  size_t nw = (size_t)(v.vsize < (long)buf.size() ? v.vsize : (long)buf.size());
  if (fseeko(fp_, v.begin + rec * recsize_, SEEK_SET) != 0 || fwrite(buf.data(), 1, nw, fp_) != nw) return false;
  if (rec + 1 > numrecs_) {
    numrecs_ = rec + 1;
    unsigned char b[4] = {(unsigned char)(numrecs_ >> 24), (unsigned char)(numrecs_ >> 16), (unsigned char)(numrecs_ >> 8), (unsigned char)numrecs_};
    fseeko(fp_, 4, SEEK_SET);
    fwrite(b, 1, 4, fp_);
  }
  return true;
}

bool NcFile::sync() { return fp_ && fflush(fp_) == 0; }

void NcFile::close() {
  if (fp_) { fflush(fp_); fclose(fp_); fp_ = nullptr; }
}

}  // namespace cgfdm
