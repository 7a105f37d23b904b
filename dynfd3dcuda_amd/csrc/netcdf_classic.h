// Minimal netCDF classic-format (CDF-1 / CDF-2) reader and writer, written
// from the published format specification.  It covers what the solver's
// inputs and outputs use: fixed-size and record variables of type
// NC_INT / NC_FLOAT / NC_DOUBLE (NC_BYTE/NC_SHORT read only), one unlimited
// dimension, global/variable attributes skipped on read.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace cgfdm {

enum NcType { NC_BYTE_ = 1, NC_CHAR_ = 2, NC_SHORT_ = 3, NC_INT_ = 4, NC_FLOAT_ = 5, NC_DOUBLE_ = 6 };

class NcFile {
 public:
  struct Dim { std::string name; long len; };  // len 0 = record dimension
  struct Var {
    std::string name;
    std::vector<int> dims;
    int type = NC_FLOAT_;
    long vsize = 0;      // bytes of one record (record vars) or of the whole var
    long long begin = 0;
    bool is_record = false;
    long count() const;  // elements per record / total (fixed vars)
  };

  ~NcFile();
  // reading
  bool open_read(const std::string& path, std::string& err);
  bool read_float(const std::string& var, std::vector<float>& out, std::string& err);
  bool read_double(const std::string& var, std::vector<double>& out, std::string& err);
  bool read_int(const std::string& var, std::vector<int>& out, std::string& err);
  long numrecs() const { return numrecs_; }
  const std::vector<Dim>& dims() const { return dims_; }
  const std::vector<Var>& vars() const { return vars_; }
  // writing: create (NC_CLOBBER), define, end_def, then put
  bool create(const std::string& path, std::string& err);
  int def_dim(const std::string& name, long len);  // len 0 = unlimited
  int def_var(const std::string& name, int type, const std::vector<int>& dims);
  bool end_def(std::string& err);
  bool put_var(int varid, const void* data);                 // fixed-size variable
  bool put_record(int varid, long rec, const void* data);    // record variable
  bool sync();
  void close();

 private:
  bool read_header(std::string& err);
  const Var* find(const std::string& name) const;
  bool read_raw(const Var& v, std::vector<unsigned char>& raw, std::string& err);
  FILE* fp_ = nullptr;
  bool writing_ = false;
  int version_ = 1;
  long numrecs_ = 0;
  long long recsize_ = 0;
  std::vector<Dim> dims_;
  std::vector<Var> vars_;
};

}  // namespace cgfdm
