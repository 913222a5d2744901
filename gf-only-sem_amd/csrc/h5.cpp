// h5.cpp -- HDF5 1.10 C API bound at run time (see h5.hpp).
#include "h5.hpp"

#include <dlfcn.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <sstream>

namespace gfx::h5 {
namespace {

// Values of the HDF5 1.10 public headers (H5Fpublic.h, H5Spublic.h,
// H5public.h, H5Tpublic.h) used below.
constexpr unsigned ACC_RDONLY = 0x0000u, ACC_RDWR = 0x0001u, ACC_TRUNC = 0x0002u;
constexpr hid_t P_DEFAULT = 0, S_ALL = 0, E_DEFAULT = 0;
constexpr int S_SCALAR = 0;
constexpr int INDEX_NAME = 0, ITER_INC = 0, ITER_NATIVE = 2;
constexpr int T_CLASS_INTEGER = 0, T_CLASS_FLOAT = 1, T_CLASS_STRING = 3;

struct GInfo {  // H5G_info_t
    int storage_type;
    hsize_t nlinks;
    int64_t max_corder;
    bool mounted;
};

struct Api {
    void* so = nullptr;
    std::string why;
    int (*open)();
    herr_t (*eset_auto2)(hid_t, void*, void*);
    hid_t (*fopen)(const char*, unsigned, hid_t);
    hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*gopen2)(hid_t, const char*, hid_t);
    hid_t (*gcreate2)(hid_t, const char*, hid_t, hid_t, hid_t);
    herr_t (*gclose)(hid_t);
    herr_t (*gget_info)(hid_t, GInfo*);
    hid_t (*dopen2)(hid_t, const char*, hid_t);
    hid_t (*dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
    herr_t (*dclose)(hid_t);
    hid_t (*dget_space)(hid_t);
    hid_t (*dget_type)(hid_t);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
    herr_t (*sclose)(hid_t);
    hid_t (*screate)(int);
    hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
    int (*sget_ndims)(hid_t);
    int (*sget_dims)(hid_t, hsize_t*, hsize_t*);
    hid_t (*pcreate)(hid_t);
    herr_t (*pset_chunk)(hid_t, int, const hsize_t*);
    herr_t (*pset_deflate)(hid_t, unsigned);
    herr_t (*pset_shuffle)(hid_t);
    herr_t (*sselect_hyperslab)(hid_t, int, const hsize_t*, const hsize_t*, const hsize_t*, const hsize_t*);
    herr_t (*pclose)(hid_t);
    htri_t (*lexists)(hid_t, const char*, hid_t);
    herr_t (*ldelete)(hid_t, const char*, hid_t);
    long (*lget_name_by_idx)(hid_t, const char*, int, int, hsize_t, char*, size_t, hid_t);
    htri_t (*aexists)(hid_t, const char*);
    hid_t (*aopen)(hid_t, const char*, hid_t);
    herr_t (*aread)(hid_t, hid_t, void*);
    hid_t (*acreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t);
    herr_t (*awrite)(hid_t, hid_t, const void*);
    herr_t (*aclose)(hid_t);
    herr_t (*adelete)(hid_t, const char*);
    hid_t (*aget_type)(hid_t);
    hid_t (*aget_space)(hid_t);
    long (*aget_name_by_idx)(hid_t, const char*, int, int, hsize_t, char*, size_t, hid_t);
    hid_t (*tcopy)(hid_t);
    herr_t (*tset_size)(hid_t, size_t);
    herr_t (*tclose)(hid_t);
    size_t (*tget_size)(hid_t);
    htri_t (*tis_variable_str)(hid_t);
    int (*tget_class)(hid_t);
    herr_t (*free_memory)(void*);
    hid_t *g_double, *g_float, *g_int, *g_int64, *g_hbool, *g_uchar, *g_c_s1, *g_dcpl;
};

std::recursive_mutex g_mu;

template <typename F>
bool bind(void* so, const char* name, F& fn, std::string& why) {
    void* p = dlsym(so, name);
    if (!p) {
        why = std::string("symbol ") + name + " missing in the HDF5 library";
        return false;
    }
    fn = reinterpret_cast<F>(p);
    return true;
}

Api g_api;
bool g_tried = false;

Api* load() {
    Api& api = g_api;
    if (g_tried) return api.so ? &api : nullptr;
    g_tried = true;
    std::vector<std::string> cand;
    if (const char* e = std::getenv("GFX_HDF5_LIB")) cand.emplace_back(e);
    cand.insert(cand.end(), {"libhdf5.so.103", "libhdf5_serial.so.103", "/opt/conda/lib/libhdf5.so.103",
                             "libhdf5.so"});
    for (const auto& c : cand) {
        api.so = dlopen(c.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (api.so) break;
    }
    if (!api.so) {
        api.why = "no HDF5 1.10 library found (set GFX_HDF5_LIB to libhdf5.so.103)";
        return nullptr;
    }
    std::string& w = api.why;
    bool ok = bind(api.so, "H5open", api.open, w) && bind(api.so, "H5Eset_auto2", api.eset_auto2, w) &&
              bind(api.so, "H5Fopen", api.fopen, w) && bind(api.so, "H5Fcreate", api.fcreate, w) &&
              bind(api.so, "H5Fclose", api.fclose, w) && bind(api.so, "H5Gopen2", api.gopen2, w) &&
              bind(api.so, "H5Gcreate2", api.gcreate2, w) && bind(api.so, "H5Gclose", api.gclose, w) &&
              bind(api.so, "H5Gget_info", api.gget_info, w) && bind(api.so, "H5Dopen2", api.dopen2, w) &&
              bind(api.so, "H5Dcreate2", api.dcreate2, w) && bind(api.so, "H5Dclose", api.dclose, w) &&
              bind(api.so, "H5Dget_space", api.dget_space, w) &&
              bind(api.so, "H5Dget_type", api.dget_type, w) && bind(api.so, "H5Dread", api.dread, w) &&
              bind(api.so, "H5Dwrite", api.dwrite, w) && bind(api.so, "H5Sclose", api.sclose, w) &&
              bind(api.so, "H5Screate", api.screate, w) &&
              bind(api.so, "H5Screate_simple", api.screate_simple, w) &&
              bind(api.so, "H5Sget_simple_extent_ndims", api.sget_ndims, w) &&
              bind(api.so, "H5Sget_simple_extent_dims", api.sget_dims, w) &&
              bind(api.so, "H5Pcreate", api.pcreate, w) && bind(api.so, "H5Pset_chunk", api.pset_chunk, w) &&
              bind(api.so, "H5Pset_deflate", api.pset_deflate, w) && bind(api.so, "H5Pclose", api.pclose, w) &&
              bind(api.so, "H5Pset_shuffle", api.pset_shuffle, w) &&
              bind(api.so, "H5Sselect_hyperslab", api.sselect_hyperslab, w) &&
              bind(api.so, "H5Lexists", api.lexists, w) && bind(api.so, "H5Ldelete", api.ldelete, w) &&
              bind(api.so, "H5Lget_name_by_idx", api.lget_name_by_idx, w) &&
              bind(api.so, "H5Aexists", api.aexists, w) && bind(api.so, "H5Aopen", api.aopen, w) &&
              bind(api.so, "H5Aread", api.aread, w) && bind(api.so, "H5Acreate2", api.acreate2, w) &&
              bind(api.so, "H5Awrite", api.awrite, w) && bind(api.so, "H5Aclose", api.aclose, w) &&
              bind(api.so, "H5Adelete", api.adelete, w) && bind(api.so, "H5Aget_type", api.aget_type, w) &&
              bind(api.so, "H5Aget_space", api.aget_space, w) &&
              bind(api.so, "H5Aget_name_by_idx", api.aget_name_by_idx, w) &&
              bind(api.so, "H5Tcopy", api.tcopy, w) && bind(api.so, "H5Tset_size", api.tset_size, w) &&
              bind(api.so, "H5Tclose", api.tclose, w) && bind(api.so, "H5Tget_size", api.tget_size, w) &&
              bind(api.so, "H5Tis_variable_str", api.tis_variable_str, w) &&
              bind(api.so, "H5Tget_class", api.tget_class, w) &&
              bind(api.so, "H5free_memory", api.free_memory, w) &&
              bind(api.so, "H5T_NATIVE_DOUBLE_g", api.g_double, w) &&
              bind(api.so, "H5T_NATIVE_FLOAT_g", api.g_float, w) &&
              bind(api.so, "H5T_NATIVE_INT_g", api.g_int, w) &&
              bind(api.so, "H5T_NATIVE_INT64_g", api.g_int64, w) &&
              bind(api.so, "H5T_NATIVE_HBOOL_g", api.g_hbool, w) && bind(api.so, "H5T_C_S1_g", api.g_c_s1, w) &&
              bind(api.so, "H5T_NATIVE_UCHAR_g", api.g_uchar, w) &&
              bind(api.so, "H5P_CLS_DATASET_CREATE_ID_g", api.g_dcpl, w);
    if (!ok || api.open() < 0) {
        if (ok) w = "H5open failed";
        dlclose(api.so);
        api.so = nullptr;
        return nullptr;
    }
    api.eset_auto2(E_DEFAULT, nullptr, nullptr);  // errors come back as exceptions instead
    return &api;
}

Api& A() {
    Api* a = load();
    if (!a) throw Error(GFX_ERR_IO, "HDF5 unavailable: " + g_api.why);
    return *a;
}

[[noreturn]] void fail(const std::string& what) { throw Error(GFX_ERR_IO, what); }

hid_t native(T t) {
    Api& a = A();
    switch (t) {
        case T::F64: return *a.g_double;
        case T::F32: return *a.g_float;
        case T::I32: return *a.g_int;
        case T::I64: return *a.g_int64;
        case T::HBOOL: return *a.g_hbool;
        case T::U8: return *a.g_uchar;
    }
    return *a.g_double;
}

template <typename V>
T type_of();
template <>
T type_of<double>() { return T::F64; }
template <>
T type_of<float>() { return T::F32; }
template <>
T type_of<int32_t>() { return T::I32; }
template <>
T type_of<int64_t>() { return T::I64; }
template <>
T type_of<uint8_t>() { return T::U8; }

}  // namespace

Obj& Obj::operator=(Obj&& o) noexcept {
    if (this != &o) {
        close();
        id_ = o.id_;
        kind_ = o.kind_;
        o.id_ = -1;
    }
    return *this;
}

void Obj::close() {
    if (id_ < 0) return;
    Api* a = load();
    if (a) {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        switch (kind_) {
            case FILE: a->fclose(id_); break;
            case GROUP: a->gclose(id_); break;
            case DSET: a->dclose(id_); break;
            case SPACE: a->sclose(id_); break;
            case ATTR: a->aclose(id_); break;
            case TYPE: a->tclose(id_); break;
            case PLIST: a->pclose(id_); break;
        }
    }
    id_ = -1;
}

bool available(std::string* why) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api* a = load();
    if (!a && why) *why = g_api.why;
    return a != nullptr;
}

Obj open_read(const std::string& path) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    hid_t f = A().fopen(path.c_str(), ACC_RDONLY, P_DEFAULT);
    if (f < 0) fail("H5Fopen failed: " + path);
    return Obj(f, Obj::FILE);
}

Obj open_rw(const std::string& path) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    hid_t f = A().fopen(path.c_str(), ACC_RDWR, P_DEFAULT);
    if (f < 0) fail("H5Fopen (read-write) failed: " + path);
    return Obj(f, Obj::FILE);
}

Obj create_trunc(const std::string& path) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    hid_t f = A().fcreate(path.c_str(), ACC_TRUNC, P_DEFAULT, P_DEFAULT);
    if (f < 0) fail("H5Fcreate failed: " + path);
    return Obj(f, Obj::FILE);
}

bool exists(hid_t loc, const std::string& path) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    // walk the path so a missing intermediate group reads as "absent"
    std::string cur;
    size_t i = 0;
    if (!path.empty() && path[0] == '/') {
        cur = "/";
        i = 1;
    }
    while (i <= path.size()) {
        size_t j = path.find('/', i);
        if (j == std::string::npos) j = path.size();
        if (j > i) {
            cur += (cur.empty() || cur.back() == '/') ? path.substr(i, j - i) : "/" + path.substr(i, j - i);
            if (a.lexists(loc, cur.c_str(), P_DEFAULT) <= 0) return false;
        }
        i = j + 1;
    }
    return true;
}

bool attr_exists(hid_t loc, const std::string& name) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    return A().aexists(loc, name.c_str()) > 0;
}

Obj open_group(hid_t loc, const std::string& path) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    hid_t g = A().gopen2(loc, path.c_str(), P_DEFAULT);
    if (g < 0) fail("H5Gopen2 failed for: " + path);
    return Obj(g, Obj::GROUP);
}

Obj open_object(hid_t loc, const std::string& path) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    hid_t g = a.gopen2(loc, path.c_str(), P_DEFAULT);
    if (g >= 0) return Obj(g, Obj::GROUP);
    hid_t d = a.dopen2(loc, path.c_str(), P_DEFAULT);
    if (d >= 0) return Obj(d, Obj::DSET);
    fail("no group or dataset: " + path);
}

Obj require_group(hid_t loc, const std::string& path) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    std::string cur;
    size_t i = 0;
    if (!path.empty() && path[0] == '/') {
        cur = "/";
        i = 1;
    }
    while (i <= path.size()) {
        size_t j = path.find('/', i);
        if (j == std::string::npos) j = path.size();
        if (j > i) {
            cur += (cur.empty() || cur.back() == '/') ? path.substr(i, j - i) : "/" + path.substr(i, j - i);
            if (a.lexists(loc, cur.c_str(), P_DEFAULT) <= 0) {
                hid_t g = a.gcreate2(loc, cur.c_str(), P_DEFAULT, P_DEFAULT, P_DEFAULT);
                if (g < 0) fail("H5Gcreate2 failed for: " + cur);
                a.gclose(g);
            }
        }
        i = j + 1;
    }
    return open_group(loc, path.empty() ? "/" : path);
}

std::vector<hsize_t> shape(hid_t loc, const std::string& path) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    Obj d(a.dopen2(loc, path.c_str(), P_DEFAULT), Obj::DSET);
    if (!d.valid()) fail("H5Dopen2 failed for: " + path);
    Obj s(a.dget_space(d.id()), Obj::SPACE);
    if (!s.valid()) fail("H5Dget_space failed for: " + path);
    const int nd = a.sget_ndims(s.id());
    if (nd < 0 || nd > 16) fail("bad rank for: " + path);
    std::vector<hsize_t> dims(nd);
    if (nd > 0) a.sget_dims(s.id(), dims.data(), nullptr);
    return dims;
}

template <typename V>
std::vector<V> read(hid_t loc, const std::string& path, std::vector<hsize_t>* dims_out) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    Obj d(a.dopen2(loc, path.c_str(), P_DEFAULT), Obj::DSET);
    if (!d.valid()) fail("H5Dopen2 failed for: " + path);
    Obj s(a.dget_space(d.id()), Obj::SPACE);
    if (!s.valid()) fail("H5Dget_space failed for: " + path);
    const int nd = a.sget_ndims(s.id());
    if (nd < 0 || nd > 16) fail("H5Sget_simple_extent_ndims failed for: " + path);
    std::vector<hsize_t> dims(nd);
    if (nd > 0) a.sget_dims(s.id(), dims.data(), nullptr);
    size_t total = 1;
    for (auto v : dims) total *= size_t(v);
    std::vector<V> out(total);
    if (total > 0 && a.dread(d.id(), native(type_of<V>()), S_ALL, S_ALL, P_DEFAULT, out.data()) < 0)
        fail("H5Dread failed for: " + path);
    if (dims_out) *dims_out = dims;
    return out;
}

template std::vector<double> read<double>(hid_t, const std::string&, std::vector<hsize_t>*);
template std::vector<float> read<float>(hid_t, const std::string&, std::vector<hsize_t>*);
template std::vector<int32_t> read<int32_t>(hid_t, const std::string&, std::vector<hsize_t>*);
template std::vector<int64_t> read<int64_t>(hid_t, const std::string&, std::vector<hsize_t>*);
template std::vector<uint8_t> read<uint8_t>(hid_t, const std::string&, std::vector<hsize_t>*);

void write(hid_t loc, const std::string& path, T file_type, T mem_type, const void* data,
           const std::vector<hsize_t>& dims, const WriteOpts& o) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    if (o.overwrite && exists(loc, path)) a.ldelete(loc, path.c_str(), P_DEFAULT);
    // parent groups
    const size_t slash = path.find_last_of('/');
    if (slash != std::string::npos && slash > 0) require_group(loc, path.substr(0, slash));
    Obj space(a.screate_simple(int(dims.size()), dims.data(), nullptr), Obj::SPACE);
    if (!space.valid()) fail("H5Screate_simple failed for " + path);
    size_t total = 1;
    for (auto v : dims) total *= size_t(v);
    Obj plist(a.pcreate(*a.g_dcpl), Obj::PLIST);
    if (!plist.valid()) fail("H5Pcreate failed for " + path);
    if ((o.chunk_whole || o.deflate > 0) && total > 0) {
        if (a.pset_chunk(plist.id(), int(dims.size()), dims.data()) < 0) fail("H5Pset_chunk failed for " + path);
        if (o.deflate > 0 && a.pset_deflate(plist.id(), unsigned(o.deflate)) < 0)
            fail("H5Pset_deflate failed for " + path);
    }
    Obj d(a.dcreate2(loc, path.c_str(), native(file_type), space.id(), P_DEFAULT, plist.id(), P_DEFAULT),
          Obj::DSET);
    if (!d.valid()) fail("H5Dcreate2 failed for " + path);
    if (total == 0) return;
    // FP64 -> FP32 files are converted here with a plain cast, as the
    // reference's record writer does (record.cpp:91-95)
    if (file_type == T::F32 && mem_type == T::F64) {
        std::vector<float> buf(total);
        const double* src = static_cast<const double*>(data);
        for (size_t i = 0; i < total; ++i) buf[i] = static_cast<float>(src[i]);
        if (a.dwrite(d.id(), native(T::F32), S_ALL, S_ALL, P_DEFAULT, buf.data()) < 0)
            fail("H5Dwrite failed for " + path);
        return;
    }
    if (a.dwrite(d.id(), native(mem_type), S_ALL, S_ALL, P_DEFAULT, data) < 0) fail("H5Dwrite failed for " + path);
}

template <typename V>
void write_vec(hid_t loc, const std::string& path, const std::vector<V>& v, std::vector<hsize_t> dims,
               const WriteOpts& o) {
    if (dims.empty()) dims = {hsize_t(v.size())};
    write(loc, path, type_of<V>(), type_of<V>(), v.data(), dims, o);
}

template void write_vec<double>(hid_t, const std::string&, const std::vector<double>&, std::vector<hsize_t>,
                                const WriteOpts&);
template void write_vec<float>(hid_t, const std::string&, const std::vector<float>&, std::vector<hsize_t>,
                               const WriteOpts&);
template void write_vec<int32_t>(hid_t, const std::string&, const std::vector<int32_t>&, std::vector<hsize_t>,
                                 const WriteOpts&);
template void write_vec<int64_t>(hid_t, const std::string&, const std::vector<int64_t>&, std::vector<hsize_t>,
                                 const WriteOpts&);
template void write_vec<uint8_t>(hid_t, const std::string&, const std::vector<uint8_t>&, std::vector<hsize_t>,
                                 const WriteOpts&);

namespace {
// Number of elements of an attribute (1 for a scalar).
hsize_t attr_points(Api& a, hid_t at) {
    Obj sp(a.aget_space(at), Obj::SPACE);
    if (!sp.valid()) return 0;
    const int nd = a.sget_ndims(sp.id());
    if (nd <= 0) return nd == 0 ? 1 : 0;
    std::vector<hsize_t> d(nd);
    a.sget_dims(sp.id(), d.data(), nullptr);
    hsize_t n = 1;
    for (hsize_t v : d) n *= v;
    return n;
}
}  // namespace

bool read_attr_int(hid_t loc, const std::string& name, int& out) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    if (a.aexists(loc, name.c_str()) <= 0) return false;
    Obj at(a.aopen(loc, name.c_str(), P_DEFAULT), Obj::ATTR);
    if (!at.valid() || attr_points(a, at.id()) != 1) return false;
    return a.aread(at.id(), *a.g_int, &out) >= 0;
}

bool read_attr_double(hid_t loc, const std::string& name, double& out) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    if (a.aexists(loc, name.c_str()) <= 0) return false;
    Obj at(a.aopen(loc, name.c_str(), P_DEFAULT), Obj::ATTR);
    if (!at.valid() || attr_points(a, at.id()) != 1) return false;
    return a.aread(at.id(), *a.g_double, &out) >= 0;
}

bool read_attr_doubles(hid_t loc, const std::string& name, std::vector<double>& out) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    if (a.aexists(loc, name.c_str()) <= 0) return false;
    Obj at(a.aopen(loc, name.c_str(), P_DEFAULT), Obj::ATTR);
    if (!at.valid()) return false;
    out.assign(size_t(attr_points(a, at.id())), 0.0);
    return out.empty() || a.aread(at.id(), *a.g_double, out.data()) >= 0;
}

bool read_attr_string(hid_t loc, const std::string& name, std::string& out) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    if (a.aexists(loc, name.c_str()) <= 0) return false;
    Obj at(a.aopen(loc, name.c_str(), P_DEFAULT), Obj::ATTR);
    if (!at.valid()) return false;
    Obj ty(a.aget_type(at.id()), Obj::TYPE);
    if (!ty.valid() || a.tget_class(ty.id()) != T_CLASS_STRING) return false;
    if (a.tis_variable_str(ty.id()) > 0) {
        char* v = nullptr;
        if (a.aread(at.id(), ty.id(), &v) < 0) return false;
        out = v ? std::string(v) : std::string();
        if (v) a.free_memory(v);
        return true;
    }
    const size_t n = a.tget_size(ty.id());
    std::vector<char> buf(n + 1, '\0');
    if (a.aread(at.id(), ty.id(), buf.data()) < 0) return false;
    out = std::string(buf.data());
    return true;
}

void write_attr(hid_t loc, const std::string& name, T type, const void* value, bool overwrite) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    if (overwrite && a.aexists(loc, name.c_str()) > 0) a.adelete(loc, name.c_str());
    Obj sp(a.screate(S_SCALAR), Obj::SPACE);
    if (!sp.valid()) fail("H5Screate failed for attr: " + name);
    Obj at(a.acreate2(loc, name.c_str(), native(type), sp.id(), P_DEFAULT, P_DEFAULT), Obj::ATTR);
    if (!at.valid()) fail("H5Acreate2 failed for attr: " + name);
    if (a.awrite(at.id(), native(type), value) < 0) fail("H5Awrite failed for attr: " + name);
}

void write_attr_string(hid_t loc, const std::string& name, const std::string& value, bool overwrite,
                       size_t fixed_size) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    if (overwrite && a.aexists(loc, name.c_str()) > 0) a.adelete(loc, name.c_str());
    // fixed-length C string of the value's size (record.cpp:37-57), or of
    // fixed_size bytes, zero padded (writer.hh:131-141 uses 64)
    const size_t size = fixed_size ? fixed_size : (value.empty() ? 1 : value.size());
    if (value.size() > size) fail("string attr longer than its fixed size: " + name);
    Obj ty(a.tcopy(*a.g_c_s1), Obj::TYPE);
    if (!ty.valid() || a.tset_size(ty.id(), size) < 0) fail("string type failed for attr: " + name);
    Obj sp(a.screate(S_SCALAR), Obj::SPACE);
    Obj at(a.acreate2(loc, name.c_str(), ty.id(), sp.id(), P_DEFAULT, P_DEFAULT), Obj::ATTR);
    if (!at.valid()) fail("H5Acreate2 failed for string attr: " + name);
    std::string v = value;
    v.resize(size, '\0');
    if (a.awrite(at.id(), ty.id(), v.c_str()) < 0) fail("H5Awrite failed for string attr: " + name);
}

void write_attr_array(hid_t loc, const std::string& name, T type, const void* values, hsize_t n, bool overwrite) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    if (overwrite && a.aexists(loc, name.c_str()) > 0) a.adelete(loc, name.c_str());
    Obj sp(a.screate_simple(1, &n, nullptr), Obj::SPACE);
    if (!sp.valid()) fail("H5Screate_simple failed for attr: " + name);
    Obj at(a.acreate2(loc, name.c_str(), native(type), sp.id(), P_DEFAULT, P_DEFAULT), Obj::ATTR);
    if (!at.valid()) fail("H5Acreate2 failed for attr: " + name);
    if (a.awrite(at.id(), native(type), values) < 0) fail("H5Awrite failed for attr: " + name);
}

Obj create_dataset(hid_t loc, const std::string& path, T file_type, const std::vector<hsize_t>& dims,
                   const std::vector<hsize_t>& chunk, bool shuffle, int deflate) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    const size_t slash = path.find_last_of('/');
    if (slash != std::string::npos && slash > 0) require_group(loc, path.substr(0, slash));
    Obj space(a.screate_simple(int(dims.size()), dims.data(), nullptr), Obj::SPACE);
    if (!space.valid()) fail("H5Screate_simple failed for " + path);
    Obj plist(a.pcreate(*a.g_dcpl), Obj::PLIST);
    if (!plist.valid()) fail("H5Pcreate failed for " + path);
    if (!chunk.empty()) {
        if (chunk.size() != dims.size()) fail("chunk rank differs from dataset rank: " + path);
        if (shuffle && a.pset_shuffle(plist.id()) < 0) fail("H5Pset_shuffle failed for " + path);
        if (a.pset_chunk(plist.id(), int(chunk.size()), chunk.data()) < 0) fail("H5Pset_chunk failed for " + path);
        if (deflate > 0 && a.pset_deflate(plist.id(), unsigned(deflate)) < 0) fail("H5Pset_deflate failed for " + path);
    }
    Obj d(a.dcreate2(loc, path.c_str(), native(file_type), space.id(), P_DEFAULT, plist.id(), P_DEFAULT),
          Obj::DSET);
    if (!d.valid()) fail("H5Dcreate2 failed for " + path);
    return d;
}

void write_slab(hid_t dset, T mem_type, const void* data, const std::vector<hsize_t>& start,
                const std::vector<hsize_t>& count) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    Obj fs(a.dget_space(dset), Obj::SPACE);
    if (!fs.valid()) fail("H5Dget_space failed");
    if (a.sselect_hyperslab(fs.id(), 0 /* H5S_SELECT_SET */, start.data(), nullptr, count.data(), nullptr) < 0)
        fail("H5Sselect_hyperslab failed");
    Obj ms(a.screate_simple(int(count.size()), count.data(), nullptr), Obj::SPACE);
    if (!ms.valid()) fail("H5Screate_simple failed for a slab");
    if (a.dwrite(dset, native(mem_type), ms.id(), fs.id(), P_DEFAULT, data) < 0) fail("H5Dwrite of a slab failed");
}

std::vector<std::string> list(hid_t loc, const std::string& group) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    Obj g = open_group(loc, group);
    GInfo info{};
    if (a.gget_info(g.id(), &info) < 0) fail("H5Gget_info failed for: " + group);
    std::vector<std::string> names;
    for (hsize_t i = 0; i < info.nlinks; ++i) {
        char buf[512];
        const long n = a.lget_name_by_idx(g.id(), ".", INDEX_NAME, ITER_INC, i, buf, sizeof(buf), P_DEFAULT);
        if (n > 0) names.emplace_back(buf, size_t(n));
    }
    return names;
}

namespace {

std::string type_text(Api& a, hid_t ty) {
    const int cls = a.tget_class(ty);
    const size_t sz = a.tget_size(ty);
    std::ostringstream os;
    if (cls == T_CLASS_FLOAT) os << "f" << sz * 8;
    else if (cls == T_CLASS_INTEGER) os << "i" << sz * 8;
    else if (cls == T_CLASS_STRING) os << "str";
    else os << "class" << cls << "/" << sz;
    return os.str();
}

std::string dims_text(Api& a, hid_t space) {
    const int nd = a.sget_ndims(space);
    std::ostringstream os;
    os << "[";
    if (nd > 0) {
        std::vector<hsize_t> d(nd);
        a.sget_dims(space, d.data(), nullptr);
        for (int i = 0; i < nd; ++i) os << (i ? "," : "") << d[i];
    }
    os << "]";
    return os.str();
}

void describe_attrs(Api& a, hid_t obj, const std::string& where, std::ostringstream& os) {
    std::vector<std::string> names;
    for (hsize_t i = 0;; ++i) {
        char buf[256];
        const long n = a.aget_name_by_idx(obj, ".", INDEX_NAME, ITER_INC, i, buf, sizeof(buf), P_DEFAULT);
        if (n <= 0) break;
        names.emplace_back(buf, size_t(n));
    }
    for (const auto& n : names) {
        Obj at(a.aopen(obj, n.c_str(), P_DEFAULT), Obj::ATTR);
        if (!at.valid()) continue;
        Obj ty(a.aget_type(at.id()), Obj::TYPE);
        Obj sp(a.aget_space(at.id()), Obj::SPACE);
        os << where << "@" << n << " " << type_text(a, ty.id()) << " " << dims_text(a, sp.id()) << "\n";
    }
}

void describe_group(Api& a, hid_t file, const std::string& path, std::ostringstream& os) {
    Obj g(a.gopen2(file, path.c_str(), P_DEFAULT), Obj::GROUP);
    if (!g.valid()) return;
    describe_attrs(a, g.id(), path, os);
    GInfo info{};
    if (a.gget_info(g.id(), &info) < 0) return;
    for (hsize_t i = 0; i < info.nlinks; ++i) {
        char buf[512];
        const long n = a.lget_name_by_idx(g.id(), ".", INDEX_NAME, ITER_INC, i, buf, sizeof(buf), P_DEFAULT);
        if (n <= 0) continue;
        const std::string child = (path == "/" ? "/" : path + "/") + std::string(buf, size_t(n));
        hid_t gg = a.gopen2(file, child.c_str(), P_DEFAULT);
        if (gg >= 0) {
            a.gclose(gg);
            describe_group(a, file, child, os);
            continue;
        }
        Obj d(a.dopen2(file, child.c_str(), P_DEFAULT), Obj::DSET);
        if (!d.valid()) continue;
        Obj ty(a.dget_type(d.id()), Obj::TYPE);
        Obj sp(a.dget_space(d.id()), Obj::SPACE);
        os << child << " " << type_text(a, ty.id()) << " " << dims_text(a, sp.id()) << "\n";
        describe_attrs(a, d.id(), child, os);
    }
}

}  // namespace

std::string describe(const std::string& path) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    Api& a = A();
    Obj f = open_read(path);
    std::ostringstream os;
    describe_group(a, f.id(), "/", os);
    return os.str();
}

}  // namespace gfx::h5
