// h5.hpp -- minimal HDF5 1.10 binding resolved at run time with dlopen.
//
// The solver library must load where no HDF5 is installed (the kernels and
// the in-memory API need none), so the file layer binds the C library on
// first use: $GFX_HDF5_LIB, then libhdf5.so.103 / libhdf5.so on the loader
// path, then the conda copy under /opt/conda/lib.  Every call that fails
// throws gfx::Error(GFX_ERR_IO) naming the object, the way the reference's
// readers throw std::runtime_error (forward/share/src/io.cpp:43-101).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace gfx::h5 {

using hid_t = int64_t;
using herr_t = int;
using htri_t = int;
using hsize_t = unsigned long long;

// Scalar types understood by the typed helpers below.
enum class T { F64, F32, I32, I64, HBOOL, U8 };

// Owning handle: closes with the matching H5*close on destruction.
class Obj {
public:
    enum Kind { FILE, GROUP, DSET, SPACE, ATTR, TYPE, PLIST };
    Obj() = default;
    Obj(hid_t id, Kind k) : id_(id), kind_(k) {}
    Obj(const Obj&) = delete;
    Obj& operator=(const Obj&) = delete;
    Obj(Obj&& o) noexcept : id_(o.id_), kind_(o.kind_) { o.id_ = -1; }
    Obj& operator=(Obj&& o) noexcept;
    ~Obj() { close(); }
    void close();
    hid_t id() const { return id_; }
    bool valid() const { return id_ >= 0; }

private:
    hid_t id_ = -1;
    Kind kind_ = FILE;
};

// Is an HDF5 library available?  (Never throws.)
bool available(std::string* why = nullptr);

Obj open_read(const std::string& path);
Obj create_trunc(const std::string& path);
Obj open_rw(const std::string& path);

// Link / attribute existence (false for missing intermediate groups).
bool exists(hid_t loc, const std::string& path);
bool attr_exists(hid_t loc, const std::string& name);
Obj open_group(hid_t loc, const std::string& path);
// A group or a dataset (for attribute access).
Obj open_object(hid_t loc, const std::string& path);
// Creates every missing group along an absolute or relative path.
Obj require_group(hid_t loc, const std::string& path);

// Dataset shape (throws when missing).
std::vector<hsize_t> shape(hid_t loc, const std::string& path);

// Whole-dataset reads converted by HDF5 to the requested memory type.
template <typename V>
std::vector<V> read(hid_t loc, const std::string& path, std::vector<hsize_t>* dims = nullptr);
// Empty vector when the link does not exist (io.cpp:97-103 try_read_dataset).
template <typename V>
std::vector<V> try_read(hid_t loc, const std::string& path) {
    if (!exists(loc, path)) return {};
    return read<V>(loc, path);
}

// Dataset writes; chunk/deflate follow the reference writers (record.cpp
// writes one whole chunk per field; partitions are gzip-4 compressed).
struct WriteOpts {
    bool chunk_whole = false;
    int deflate = 0;  // 0 = none, 1..9 gzip level (implies chunking)
    bool overwrite = false;
};
void write(hid_t loc, const std::string& path, T file_type, T mem_type, const void* data,
           const std::vector<hsize_t>& dims, const WriteOpts& o = {});
template <typename V>
void write_vec(hid_t loc, const std::string& path, const std::vector<V>& v,
               std::vector<hsize_t> dims = {}, const WriteOpts& o = {});

// Scalar attributes (create, or overwrite when `overwrite`); the scalar
// readers return false for an attribute of more than one element.
bool read_attr_int(hid_t loc, const std::string& name, int& out);
bool read_attr_double(hid_t loc, const std::string& name, double& out);
bool read_attr_string(hid_t loc, const std::string& name, std::string& out);
// Every element of a numeric attribute (scalar or 1-D array) as FP64.
bool read_attr_doubles(hid_t loc, const std::string& name, std::vector<double>& out);
void write_attr(hid_t loc, const std::string& name, T type, const void* value, bool overwrite = false);
// fixed_size = 0: the value's own length; else that many bytes, zero padded.
void write_attr_string(hid_t loc, const std::string& name, const std::string& value,
                       bool overwrite = false, size_t fixed_size = 0);
// 1-D array attribute of n values.
void write_attr_array(hid_t loc, const std::string& name, T type, const void* values, hsize_t n,
                      bool overwrite = false);

// Dataset created empty (chunked when chunk is given, with optional shuffle
// and gzip level) for hyperslab writes; parent groups are created.
Obj create_dataset(hid_t loc, const std::string& path, T file_type, const std::vector<hsize_t>& dims,
                   const std::vector<hsize_t>& chunk = {}, bool shuffle = false, int deflate = 0);
// Writes the block [start, start + count) of an open dataset from memory of mem_type.
void write_slab(hid_t dset, T mem_type, const void* data, const std::vector<hsize_t>& start,
                const std::vector<hsize_t>& count);

// Names of the links directly under a group, in native order.
std::vector<std::string> list(hid_t loc, const std::string& group);

// Text description of a file: one line per dataset ("path class size dims")
// and per attribute ("path@name ..."), depth-first in name order.
std::string describe(const std::string& path);

}  // namespace gfx::h5
