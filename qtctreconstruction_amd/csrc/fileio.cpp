// fileio.cpp — the reference's on-disk formats: NPY volumes and point clouds,
// and the sinogram stack with its JSON sidecar.
//
// NPY (FileIO.cpp:56-206): magic "\x93NUMPY", version 1.0, a little-endian
// uint16 header length, then a Python-dict header
//   {'descr': '<f4', 'fortran_order': False, 'shape': (a, b, c), }
// padded with spaces and a final '\n' so that 10 + header length is a
// multiple of 16, then the raw little-endian floats. The writer produces
// those bytes exactly; the reader accepts what the reference's reader
// accepts: versions 1.x (uint16 length) and 2.x+ (uint32 length), any header
// that names '<f4', and the first three (volume) or two (point cloud) numbers
// between the first '(' and the next ')' as the shape.
//
// Sinogram files (CtReconstructionController.cpp:1353-1482 writes them,
// :1494-1616 reads them): an NPY of shape (nz, na, bins) and, beside it, the
// same path with ".npy" replaced by ".meta.json": an object with version,
// depth, num_angles, detector_bins, detector_spacing_mm, original_min_hu,
// original_max_hu and angles_deg, keys in sorted order and indented by four
// spaces, numbers in their shortest round-trip form. On reading, a missing or
// unparsable sidecar, or an angle list whose length is not na, falls back to
// the reader's defaults (spacing 1, uniform angles 180/na*a, min/max of the
// data), key by key as the reference does.
#include "ctx.h"

#include <cctype>
#include <cfloat>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>

using namespace cthost;

namespace {

struct IoError : StatusError {
    explicit IoError(const std::string& m) : StatusError(CT_ERR_IO, m) {}
};

// ------------------------------------------------------------------ NPY

std::string npy_header(const std::string& shape) {
    std::string h = "{'descr': '<f4', 'fortran_order': False, 'shape': (" + shape + "), }";
    size_t padded = h.size() + 1;
    while ((10 + padded) % 16 != 0) ++padded;
    h.append(padded - h.size() - 1, ' ');
    h.push_back('\n');
    return h;
}

void write_npy(const std::string& path, const std::string& shape, const float* data, size_t n) {
    std::ofstream out(path, std::ios::binary);
    if (!out.is_open()) throw IoError("cannot open " + path + " for writing");
    const std::string h = npy_header(shape);
    const unsigned char pre[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    out.write(reinterpret_cast<const char*>(pre), 8);
    const unsigned char len[2] = {(unsigned char)(h.size() & 0xff), (unsigned char)(h.size() >> 8)};
    out.write(reinterpret_cast<const char*>(len), 2);
    out.write(h.data(), (std::streamsize)h.size());
    out.write(reinterpret_cast<const char*>(data), (std::streamsize)(n * sizeof(float)));
    if (!out.good()) throw IoError("write failed: " + path);
}

// Reads the header; leaves `in` at the first data byte.
std::string read_npy_header(std::ifstream& in, const std::string& path) {
    char magic[6] = {};
    in.read(magic, 6);
    if (!in.good() || std::string(magic, 6) != std::string("\x93NUMPY", 6))
        throw IoError("not an NPY file: " + path);
    unsigned char ver[2] = {};
    in.read(reinterpret_cast<char*>(ver), 2);
    if (!in.good()) throw IoError("truncated NPY header: " + path);
    uint32_t hl = 0;
    if (ver[0] == 1) {
        unsigned char b[2] = {};
        in.read(reinterpret_cast<char*>(b), 2);
        hl = (uint32_t)b[0] | ((uint32_t)b[1] << 8);
    } else {
        unsigned char b[4] = {};
        in.read(reinterpret_cast<char*>(b), 4);
        hl = (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
    }
    if (!in.good() || hl == 0) throw IoError("truncated NPY header: " + path);
    std::string h(hl, '\0');
    in.read(&h[0], hl);
    if (!in.good()) throw IoError("truncated NPY header: " + path);
    if (h.find("<f4") == std::string::npos) throw IoError("NPY dtype is not '<f4': " + path);
    return h;
}

// Number of elements of a shape; CT_ERR_IO when it overflows size_t.
size_t checked_count(const size_t* dims, int k, const std::string& path) {
    size_t n = 1;
    for (int i = 0; i < k; ++i) {
        if (dims[i] != 0 && n > std::numeric_limits<size_t>::max() / 4 / dims[i])
            throw IoError("NPY shape too large: " + path);
        n *= dims[i];
    }
    return n;
}

// The first k numbers between the first '(' and the following ')'.
bool parse_shape(const std::string& h, int k, size_t* dims) {
    const size_t l = h.find('(');
    const size_t r = h.find(')', l == std::string::npos ? 0 : l + 1);
    if (l == std::string::npos || r == std::string::npos || r <= l + 1) return false;
    std::string s = h.substr(l + 1, r - l - 1);
    for (char& ch : s)
        if (ch == ',') ch = ' ';
    std::istringstream iss(s);
    for (int i = 0; i < k; ++i)
        if (!(iss >> dims[i])) return false;
    return true;
}

void read_floats(std::ifstream& in, float* dst, size_t n, const std::string& path) {
    in.read(reinterpret_cast<char*>(dst), (std::streamsize)(n * sizeof(float)));
    if (!in.good()) throw IoError("NPY data shorter than its shape: " + path);
}

// ----------------------------------------------------------------- JSON

// Shortest decimal form that reads back as the same double (std::to_chars:
// independent of the C locale, which a Qt application sets from the
// environment, where a decimal comma would make invalid JSON); null for
// non-finite values, which JSON cannot hold.
std::string json_number(double v) {
    if (!std::isfinite(v)) return "null";
    char buf[64];
    const std::to_chars_result r = std::to_chars(buf, buf + sizeof buf, v);
    return std::string(buf, r.ptr);
}

struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::map<std::string, JVal> obj;
    double to_double(double dflt) const { return kind == Num ? num : dflt; }
};

// Recursive-descent reader of one JSON document (RFC 8259). Returns false on
// any syntax error.
class JParser {
public:
    explicit JParser(const std::string& s) : s_(s) {}
    bool document(JVal& v) {
        if (!value(v, 0)) return false;
        ws();
        return i_ == s_.size();
    }

private:
    const std::string& s_;
    size_t i_ = 0;
    void ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
    }
    bool lit(const char* t) {
        const size_t n = std::strlen(t);
        if (s_.compare(i_, n, t) != 0) return false;
        i_ += n;
        return true;
    }
    bool value(JVal& v, int depth) {
        if (depth > 64) return false;
        ws();
        if (i_ >= s_.size()) return false;
        const char c = s_[i_];
        if (c == '{') return object(v, depth);
        if (c == '[') return array(v, depth);
        if (c == '"') {
            v.kind = JVal::Str;
            return string(v.str);
        }
        if (c == 't' && lit("true")) { v.kind = JVal::Bool; return true; }
        if (c == 'f' && lit("false")) { v.kind = JVal::Bool; return true; }
        if (c == 'n' && lit("null")) { v.kind = JVal::Null; return true; }
        return number(v);
    }
    bool number(JVal& v) {
        const size_t b = i_;
        if (i_ < s_.size() && s_[i_] == '-') ++i_;
        if (i_ >= s_.size() || !std::isdigit((unsigned char)s_[i_])) return false;
        if (s_[i_] == '0') ++i_;
        else while (i_ < s_.size() && std::isdigit((unsigned char)s_[i_])) ++i_;
        if (i_ < s_.size() && s_[i_] == '.') {
            ++i_;
            if (i_ >= s_.size() || !std::isdigit((unsigned char)s_[i_])) return false;
            while (i_ < s_.size() && std::isdigit((unsigned char)s_[i_])) ++i_;
        }
        if (i_ < s_.size() && (s_[i_] == 'e' || s_[i_] == 'E')) {
            ++i_;
            if (i_ < s_.size() && (s_[i_] == '+' || s_[i_] == '-')) ++i_;
            if (i_ >= s_.size() || !std::isdigit((unsigned char)s_[i_])) return false;
            while (i_ < s_.size() && std::isdigit((unsigned char)s_[i_])) ++i_;
        }
        v.kind = JVal::Num;
        // from_chars takes no leading '+' and reads '-' itself; the grammar
        // above has checked the form. Magnitudes past double range read as
        // +-inf, as strtod would give.
        const std::from_chars_result r = std::from_chars(s_.data() + b, s_.data() + i_, v.num);
        if (r.ec == std::errc::result_out_of_range) {
            const bool neg = s_[b] == '-';
            size_t e = s_.find_first_of("eE", b);
            const bool tiny = e != std::string::npos && e + 1 < i_ && s_[e + 1] == '-';
            v.num = tiny ? (neg ? -0.0 : 0.0)
                         : (neg ? -std::numeric_limits<double>::infinity()
                                : std::numeric_limits<double>::infinity());
        } else if (r.ec != std::errc()) {
            return false;
        }
        return true;
    }
    static void utf8(std::string& o, unsigned cp) {
        if (cp < 0x80) o.push_back((char)cp);
        else if (cp < 0x800) { o.push_back((char)(0xc0 | (cp >> 6))); o.push_back((char)(0x80 | (cp & 0x3f))); }
        else if (cp < 0x10000) {
            o.push_back((char)(0xe0 | (cp >> 12)));
            o.push_back((char)(0x80 | ((cp >> 6) & 0x3f)));
            o.push_back((char)(0x80 | (cp & 0x3f)));
        } else {
            o.push_back((char)(0xf0 | (cp >> 18)));
            o.push_back((char)(0x80 | ((cp >> 12) & 0x3f)));
            o.push_back((char)(0x80 | ((cp >> 6) & 0x3f)));
            o.push_back((char)(0x80 | (cp & 0x3f)));
        }
    }
    bool hex4(unsigned& cp) {
        if (i_ + 4 > s_.size()) return false;
        cp = 0;
        for (int k = 0; k < 4; ++k) {
            const char h = s_[i_++];
            cp <<= 4;
            if (h >= '0' && h <= '9') cp |= (unsigned)(h - '0');
            else if (h >= 'a' && h <= 'f') cp |= (unsigned)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') cp |= (unsigned)(h - 'A' + 10);
            else return false;
        }
        return true;
    }
    bool string(std::string& o) {
        ++i_;  // opening quote
        o.clear();
        while (i_ < s_.size()) {
            const char c = s_[i_++];
            if (c == '"') return true;
            if ((unsigned char)c < 0x20) return false;
            if (c != '\\') { o.push_back(c); continue; }
            if (i_ >= s_.size()) return false;
            const char e = s_[i_++];
            switch (e) {
                case '"': o.push_back('"'); break;
                case '\\': o.push_back('\\'); break;
                case '/': o.push_back('/'); break;
                case 'b': o.push_back('\b'); break;
                case 'f': o.push_back('\f'); break;
                case 'n': o.push_back('\n'); break;
                case 'r': o.push_back('\r'); break;
                case 't': o.push_back('\t'); break;
                case 'u': {
                    unsigned cp;
                    if (!hex4(cp)) return false;
                    if (cp >= 0xd800 && cp < 0xdc00 && i_ + 6 <= s_.size() && s_[i_] == '\\' &&
                        s_[i_ + 1] == 'u') {
                        i_ += 2;
                        unsigned lo;
                        if (!hex4(lo) || lo < 0xdc00 || lo >= 0xe000) return false;
                        cp = 0x10000 + ((cp - 0xd800) << 10) + (lo - 0xdc00);
                    }
                    utf8(o, cp);
                } break;
                default: return false;
            }
        }
        return false;
    }
    bool array(JVal& v, int depth) {
        ++i_;
        v.kind = JVal::Arr;
        ws();
        if (i_ < s_.size() && s_[i_] == ']') { ++i_; return true; }
        for (;;) {
            v.arr.emplace_back();
            if (!value(v.arr.back(), depth + 1)) return false;
            ws();
            if (i_ >= s_.size()) return false;
            if (s_[i_] == ',') { ++i_; continue; }
            if (s_[i_] == ']') { ++i_; return true; }
            return false;
        }
    }
    bool object(JVal& v, int depth) {
        ++i_;
        v.kind = JVal::Obj;
        ws();
        if (i_ < s_.size() && s_[i_] == '}') { ++i_; return true; }
        for (;;) {
            ws();
            if (i_ >= s_.size() || s_[i_] != '"') return false;
            std::string key;
            if (!string(key)) return false;
            ws();
            if (i_ >= s_.size() || s_[i_] != ':') return false;
            ++i_;
            JVal val;
            if (!value(val, depth + 1)) return false;
            v.obj[key] = std::move(val);  // a repeated key keeps its last value
            ws();
            if (i_ >= s_.size()) return false;
            if (s_[i_] == ',') { ++i_; continue; }
            if (s_[i_] == '}') { ++i_; return true; }
            return false;
        }
    }
};

std::string sidecar_path(const std::string& npy) {
    std::string p = npy;
    if (p.size() >= 4) {
        std::string tail = p.substr(p.size() - 4);
        for (char& ch : tail) ch = (char)std::tolower((unsigned char)ch);
        if (tail == ".npy") p.resize(p.size() - 4);
    }
    return p + ".meta.json";
}

void write_sidecar(const std::string& path, int nz, int na, int bins, float min_hu, float max_hu) {
    // angles as the exporter stores them: float step 180/na times double(a)
    const float step = 180.0f / (float)na;
    std::string j = "{\n    \"angles_deg\": [\n";
    for (int a = 0; a < na; ++a) {
        j += "        " + json_number((double)step * (double)a);
        j += (a + 1 < na) ? ",\n" : "\n";
    }
    j += "    ],\n";
    j += "    \"depth\": " + std::to_string(nz) + ",\n";
    j += "    \"detector_bins\": " + std::to_string(bins) + ",\n";
    j += "    \"detector_spacing_mm\": 1,\n";
    j += "    \"num_angles\": " + std::to_string(na) + ",\n";
    j += "    \"original_max_hu\": " + json_number((double)max_hu) + ",\n";
    j += "    \"original_min_hu\": " + json_number((double)min_hu) + ",\n";
    j += "    \"version\": 1\n}\n";
    std::ofstream out(path, std::ios::binary);
    if (!out.is_open()) throw IoError("cannot open " + path + " for writing");
    out.write(j.data(), (std::streamsize)j.size());
    if (!out.good()) throw IoError("write failed: " + path);
}

template <typename F>
int io_guarded(F&& f) {
    try {
        f();
        return CT_OK;
    } catch (const StatusError& e) {
        return e.code;
    } catch (const std::bad_alloc&) {
        return CT_ERR_NOMEM;
    } catch (const std::exception&) {
        return CT_ERR_IO;
    }
}

}  // namespace

extern "C" {

int ct_hip_save_volume_npy(const float* vol, size_t w, size_t h, size_t d, const char* path) {
    if (!vol || !path) return CT_ERR_INVALID;
    if (w == 0 || h == 0 || d == 0) return CT_ERR_INVALID;
    return io_guarded([&] {
        write_npy(path, std::to_string(d) + ", " + std::to_string(h) + ", " + std::to_string(w),
                  vol, w * h * d);
    });
}

int ct_hip_load_volume_npy(const char* path, int input_is_xyz_layout, float* vol, size_t capacity,
                           size_t* dims3) {
    if (!path || !dims3) return CT_ERR_INVALID;
    return io_guarded([&] {
        std::ifstream in(path, std::ios::binary);
        if (!in.is_open()) throw IoError(std::string("cannot open ") + path);
        const std::string h = read_npy_header(in, path);
        size_t s[3];
        if (!parse_shape(h, 3, s) || s[0] == 0 || s[1] == 0 || s[2] == 0)
            throw IoError(std::string("NPY shape is not three non-zero sizes: ") + path);
        checked_count(s, 3, path);
        // the Volume the reference builds: (z,y,x) files keep their order,
        // width = last axis; (x,y,z) files are transposed, width = first axis
        if (input_is_xyz_layout) {
            dims3[0] = s[0];
            dims3[1] = s[1];
            dims3[2] = s[2];
        } else {
            dims3[0] = s[2];
            dims3[1] = s[1];
            dims3[2] = s[0];
        }
        if (!vol) return;
        const size_t n = checked_count(s, 3, path);
        if (capacity < n) throw StatusError(CT_ERR_INVALID, "loadVolumeNPY: buffer too small");
        if (!input_is_xyz_layout) {
            read_floats(in, vol, n, path);
            return;
        }
        std::vector<float> raw(n);
        read_floats(in, raw.data(), n, path);
        const size_t sx = s[0], sy = s[1], sz = s[2];
        for (size_t z = 0; z < sz; ++z)
            for (size_t y = 0; y < sy; ++y)
                for (size_t x = 0; x < sx; ++x) vol[(z * sy + y) * sx + x] = raw[(x * sy + y) * sz + z];
    });
}

int ct_hip_save_point_cloud_npy(const float* xyzh, size_t n, const char* path) {
    if (!xyzh || !path || n == 0) return CT_ERR_INVALID;
    return io_guarded([&] { write_npy(path, std::to_string(n) + ", 4", xyzh, 4 * n); });
}

int ct_hip_load_point_cloud_npy(const char* path, float* xyzh, size_t capacity, size_t* n) {
    if (!path || !n) return CT_ERR_INVALID;
    return io_guarded([&] {
        std::ifstream in(path, std::ios::binary);
        if (!in.is_open()) throw IoError(std::string("cannot open ") + path);
        const std::string h = read_npy_header(in, path);
        size_t s[2];
        if (!parse_shape(h, 2, s) || s[0] == 0 || s[1] != 4)
            throw IoError(std::string("NPY shape is not (n, 4): ") + path);
        checked_count(s, 2, path);
        *n = s[0];
        if (!xyzh) return;
        if (capacity < s[0]) throw StatusError(CT_ERR_INVALID, "loadPointCloudNPY: buffer too small");
        read_floats(in, xyzh, 4 * s[0], path);
    });
}

int ct_hip_load_sinograms(const char* npy_path, float* sinos, size_t capacity, float* angles_deg,
                          ct_sino_info* info) {
    if (!npy_path || !info) return CT_ERR_INVALID;
    return io_guarded([&] {
        size_t dims[3];
        int rc = ct_hip_load_volume_npy(npy_path, 0, nullptr, 0, dims);
        if (rc != CT_OK) throw StatusError(rc, std::string("cannot read ") + npy_path);
        ct_sino_info r{};
        r.bins = dims[0];
        r.na = dims[1];
        r.nz = dims[2];
        r.detector_spacing_mm = 1.0f;
        r.min_hu = FLT_MAX;
        r.max_hu = -FLT_MAX;
        std::vector<float> angles;
        {
            std::ifstream jf(sidecar_path(npy_path), std::ios::binary);
            if (jf.is_open()) {
                std::stringstream ss;
                ss << jf.rdbuf();
                JVal doc;
                if (JParser(ss.str()).document(doc) && doc.kind == JVal::Obj) {
                    r.meta_found = 1;
                    const auto& o = doc.obj;
                    auto it = o.find("detector_spacing_mm");
                    if (it != o.end()) r.detector_spacing_mm = (float)it->second.to_double(1.0);
                    auto lo = o.find("original_min_hu"), hi = o.find("original_max_hu");
                    if (lo != o.end() && hi != o.end()) {
                        r.min_hu = (float)lo->second.to_double(0.0);
                        r.max_hu = (float)hi->second.to_double(0.0);
                        r.meta_minmax = 1;
                    }
                    it = o.find("angles_deg");
                    if (it != o.end() && it->second.kind == JVal::Arr)
                        for (const JVal& v : it->second.arr) angles.push_back((float)v.to_double(0.0));
                }
            }
        }
        if (angles.size() == r.na) {
            r.meta_angles = 1;
        } else {
            angles.assign(r.na, 0.0f);
            const float step = 180.0f / (float)r.na;
            for (size_t a = 0; a < r.na; ++a) angles[a] = step * (float)a;
        }
        if (sinos) {
            const size_t n = r.nz * r.na * r.bins;
            if (capacity < n) throw StatusError(CT_ERR_INVALID, "loadSinograms: buffer too small");
            rc = ct_hip_load_volume_npy(npy_path, 0, sinos, capacity, dims);
            if (rc != CT_OK) throw StatusError(rc, std::string("cannot read ") + npy_path);
            if (!r.meta_minmax)
                for (size_t i = 0; i < n; ++i) {
                    if (sinos[i] < r.min_hu) r.min_hu = sinos[i];
                    if (sinos[i] > r.max_hu) r.max_hu = sinos[i];
                }
        }
        if (angles_deg) std::memcpy(angles_deg, angles.data(), sizeof(float) * r.na);
        *info = r;
    });
}

}  // extern "C"

// Called by ct_hip_export_sinograms (capi.cpp) once the sinograms are on the host.
namespace cthost {
void write_sinogram_files(const std::string& npy_path, const float* sinos, int nz, int na,
                          int bins, float min_hu, float max_hu) {
    write_npy(npy_path,
              std::to_string(nz) + ", " + std::to_string(na) + ", " + std::to_string(bins), sinos,
              (size_t)nz * na * bins);
    write_sidecar(sidecar_path(npy_path), nz, na, bins, min_hu, max_hu);
}
}  // namespace cthost
