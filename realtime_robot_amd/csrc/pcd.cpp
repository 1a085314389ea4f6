// PCD v0.7 reader/writer (the on-disk format of every cloud the reference
// loads and saves: pcl::io::loadPCDFile at RealTimeRobot.cpp:34-35,
// savePCDFileASCII at RealTimeRobot.cpp:108-109 and function.h:126-127).
// Supports DATA ascii and binary with fields of any PCL scalar type;
// binary_compressed is rejected with RTR_ERR_FORMAT.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/rtr.h"

namespace rtr {
void set_error(const char* fmt, ...);
}
using rtr::set_error;

namespace {

struct Field {
  std::string name;
  int size = 4;
  char type = 'F';
  int count = 1;
  size_t offset = 0;  // byte offset in a binary record / column in ascii
};

struct Header {
  std::vector<Field> fields;
  size_t points = 0;
  std::string data;
  size_t data_pos = 0;
  size_t record = 0;
};

int parse_header(const std::string& buf, Header& h) {
  size_t pos = 0;
  size_t width = 0, height = 1;
  bool have_points = false;
  std::vector<int> sizes, counts;
  std::vector<char> types;
  std::vector<std::string> names;
  while (pos < buf.size()) {
    size_t end = buf.find('\n', pos);
    if (end == std::string::npos) end = buf.size();
    std::string line = buf.substr(pos, end - pos);
    pos = end + 1;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ss(line);
    std::string key;
    ss >> key;
    if (key == "FIELDS" || key == "COLUMNS") {
      std::string f;
      while (ss >> f) names.push_back(f);
    } else if (key == "SIZE") {
      int v;
      while (ss >> v) sizes.push_back(v);
    } else if (key == "TYPE") {
      std::string t;
      while (ss >> t) types.push_back(t.empty() ? 'F' : t[0]);
    } else if (key == "COUNT") {
      int v;
      while (ss >> v) counts.push_back(v);
    } else if (key == "WIDTH") {
      ss >> width;
    } else if (key == "HEIGHT") {
      ss >> height;
    } else if (key == "POINTS") {
      ss >> h.points;
      have_points = true;
    } else if (key == "DATA") {
      ss >> h.data;
      h.data_pos = pos;
      break;
    }
  }
  if (h.data.empty() || names.empty()) {
    set_error("PCD header incomplete");
    return RTR_ERR_FORMAT;
  }
  if (!have_points) {
    if (height != 0 && width > SIZE_MAX / height) {
      set_error("PCD header: WIDTH x HEIGHT overflows");
      return RTR_ERR_FORMAT;
    }
    h.points = width * height;
  }
  if (counts.empty()) counts.assign(names.size(), 1);
  if (sizes.size() != names.size() || types.size() != names.size() || counts.size() != names.size()) {
    set_error("PCD header: FIELDS/SIZE/TYPE/COUNT lengths differ");
    return RTR_ERR_FORMAT;
  }
  size_t off = 0, col = 0;
  for (size_t i = 0; i < names.size(); ++i) {
    Field f;
    f.name = names[i];
    f.size = sizes[i];
    f.type = types[i];
    f.count = counts[i];
    // PCL's scalar types: F 4 or 8 bytes, U and I 1, 2, 4 or 8 bytes
    const bool size_ok = f.type == 'F' ? (f.size == 4 || f.size == 8)
                                       : ((f.type == 'U' || f.type == 'I') &&
                                          (f.size == 1 || f.size == 2 || f.size == 4 || f.size == 8));
    if (!size_ok) {
      set_error("PCD header: field %s has TYPE %c with SIZE %d", f.name.c_str(), f.type, f.size);
      return RTR_ERR_FORMAT;
    }
    if (f.count < 1 || f.count > (1 << 20)) {
      set_error("PCD header: field %s has COUNT %d", f.name.c_str(), f.count);
      return RTR_ERR_FORMAT;
    }
    f.offset = h.data == "ascii" ? col : off;
    off += (size_t)f.size * (size_t)f.count;
    col += (size_t)f.count;
    h.fields.push_back(f);
  }
  h.record = off;
  return RTR_OK;
}

double read_scalar(const unsigned char* p, const Field& f) {
  switch (f.type) {
    case 'F':
      if (f.size == 4) {
        float v;
        memcpy(&v, p, 4);
        return v;
      } else {
        double v;
        memcpy(&v, p, 8);
        return v;
      }
    case 'U':
      if (f.size == 1) return *p;
      if (f.size == 2) {
        uint16_t v;
        memcpy(&v, p, 2);
        return v;
      }
      if (f.size == 8) {
        uint64_t v;
        memcpy(&v, p, 8);
        return (double)v;
      }
      {
        uint32_t v;
        memcpy(&v, p, 4);
        return v;
      }
    default:
      if (f.size == 1) return (int8_t)*p;
      if (f.size == 2) {
        int16_t v;
        memcpy(&v, p, 2);
        return v;
      }
      if (f.size == 8) {
        int64_t v;
        memcpy(&v, p, 8);
        return (double)v;
      }
      {
        int32_t v;
        memcpy(&v, p, 4);
        return v;
      }
  }
}

int load(const char* path, std::string& buf, Header& h) {
  if (!path) return RTR_ERR_INVALID_ARG;
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    set_error("cannot open %s", path);
    return RTR_ERR_IO;
  }
  std::ostringstream os;
  os << f.rdbuf();
  buf = os.str();
  int rc = parse_header(buf, h);
  if (rc) return rc;
  if (h.data != "ascii" && h.data != "binary") {
    set_error("PCD DATA %s is not supported", h.data.c_str());
    return RTR_ERR_FORMAT;
  }
  if (h.data == "binary") {
    // POINTS records of `record` bytes must fit the payload (checked without
    // forming points * record, which can overflow)
    const size_t avail = buf.size() >= h.data_pos ? buf.size() - h.data_pos : 0;
    if (h.record == 0 || h.points > avail / h.record) {
      set_error("PCD binary payload truncated: %zu points of %zu bytes, %zu bytes present", h.points, h.record,
                avail);
      return RTR_ERR_FORMAT;
    }
  }
  return RTR_OK;
}

int find(const Header& h, const char* name) {
  for (size_t i = 0; i < h.fields.size(); ++i)
    if (h.fields[i].name == name) return (int)i;
  return -1;
}

}  // namespace

extern "C" {

int rtr_pcd_info(const char* path, size_t* n_points, int* has_normals) {
  try {
    std::string buf;
    Header h;
    int rc = load(path, buf, h);
    if (rc) return rc;
    if (n_points) *n_points = h.points;
    if (has_normals)
      *has_normals = find(h, "normal_x") >= 0 && find(h, "normal_y") >= 0 && find(h, "normal_z") >= 0;
    return RTR_OK;
  } catch (...) {
    set_error("host allocation failed");
    return RTR_ERR_CAPACITY;
  }
}

int rtr_pcd_read(const char* path, float* xyz, float* normals, size_t cap, size_t* n_out) {
  try {
    if (!xyz || !n_out) return RTR_ERR_INVALID_ARG;
    std::string buf;
    Header h;
    int rc = load(path, buf, h);
    if (rc) return rc;
    *n_out = h.points;
    if (h.points > cap) {
      set_error("PCD has %zu points, buffer holds %zu", h.points, cap);
      return RTR_ERR_CAPACITY;
    }
    const int ix = find(h, "x"), iy = find(h, "y"), iz = find(h, "z");
    if (ix < 0 || iy < 0 || iz < 0) {
      set_error("PCD has no x/y/z fields");
      return RTR_ERR_FORMAT;
    }
    const int nxi = find(h, "normal_x"), nyi = find(h, "normal_y"), nzi = find(h, "normal_z");
    const bool has_n = nxi >= 0 && nyi >= 0 && nzi >= 0;
    const int want[6] = {ix, iy, iz, nxi, nyi, nzi};
    if (h.data == "binary") {
      const unsigned char* base = reinterpret_cast<const unsigned char*>(buf.data()) + h.data_pos;
      for (size_t i = 0; i < h.points; ++i) {
        const unsigned char* rec = base + i * h.record;
        for (int a = 0; a < 3; ++a) xyz[3 * i + a] = (float)read_scalar(rec + h.fields[want[a]].offset, h.fields[want[a]]);
        if (normals && has_n)
          for (int a = 0; a < 3; ++a)
            normals[3 * i + a] = (float)read_scalar(rec + h.fields[want[3 + a]].offset, h.fields[want[3 + a]]);
      }
    } else {
      const char* p = buf.data() + h.data_pos;
      const char* e = buf.data() + buf.size();
      size_t ncol = 0;
      for (const auto& f : h.fields) ncol += f.count;
      std::vector<double> row(ncol);
      for (size_t i = 0; i < h.points; ++i) {
        for (size_t c = 0; c < ncol; ++c) {
          while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
          if (p >= e) {
            set_error("PCD ascii payload truncated at point %zu", i);
            return RTR_ERR_FORMAT;
          }
          char* q = nullptr;
          double v = strtod(p, &q);
          if (q == p) {
            // tokens such as "nan" are parsed by strtod; anything else is bad
            set_error("PCD ascii: bad number at point %zu", i);
            return RTR_ERR_FORMAT;
          }
          row[c] = v;
          p = q;
        }
        for (int a = 0; a < 3; ++a) xyz[3 * i + a] = (float)row[h.fields[want[a]].offset];
        if (normals && has_n)
          for (int a = 0; a < 3; ++a) normals[3 * i + a] = (float)row[h.fields[want[3 + a]].offset];
      }
    }
    return RTR_OK;
  } catch (...) {
    set_error("host allocation failed");
    return RTR_ERR_CAPACITY;
  }
}

int rtr_pcd_write_ascii(const char* path, const float* xyz, size_t n, size_t stride) {
  if (!path || (n && !xyz) || (stride != 0 && stride < 12)) return RTR_ERR_INVALID_ARG;
  if (stride == 0) stride = 12;
  FILE* f = fopen(path, "wb");
  if (!f) {
    set_error("cannot open %s for writing", path);
    return RTR_ERR_IO;
  }
  fprintf(f,
          "# .PCD v0.7 - Point Cloud Data file format\nVERSION 0.7\nFIELDS x y z\nSIZE 4 4 4\n"
          "TYPE F F F\nCOUNT 1 1 1\nWIDTH %zu\nHEIGHT 1\nVIEWPOINT 0 0 0 1 0 0 0\nPOINTS %zu\n"
          "DATA ascii\n",
          n, n);
  const char* b = reinterpret_cast<const char*>(xyz);
  for (size_t i = 0; i < n; ++i) {
    float p[3];
    memcpy(p, b + i * stride, 12);
    fprintf(f, "%.8g %.8g %.8g\n", (double)p[0], (double)p[1], (double)p[2]);
  }
  if (fclose(f) != 0) {
    set_error("write to %s failed", path);
    return RTR_ERR_IO;
  }
  return RTR_OK;
}

}  // extern "C"
