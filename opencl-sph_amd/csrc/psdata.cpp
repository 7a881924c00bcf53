// psdata builder and accessors.
//
// A $psdata_specification section holds one field per line,
//     name | dims | type | default values
// and the builder lays every field out in one zero-initialised blob.  The
// layout rules are those of src/build_psdata.c, kept exactly so that blobs,
// offsets and metadata arrays are byte-identical to the reference's:
//   * lines split on '\n', '#'-lines skipped; segments split on '|', words on
//     ',' and ' ', with delimiter runs collapsing (:184-326, :585-613);
//   * a dimension word that parses (strtol, base 0) to 0 is looked up as the
//     first 4 bytes of an earlier field's default value (:29-37, :228-233);
//   * defaults are parsed only for the type strings "double", "float",
//     "unsigned int" and "int" (:255-293);
//   * gridres[3], gridcell[pnum], gridcount[rx,ry,rz], celloffset[rx,ry,rz]
//     and cellparticles[pnum] are appended, with r = (unsigned)(|hi-lo|/h)
//     evaluated in double (:40-183);
//   * fields are regrouped as double, float, "unsigned int", then the rest,
//     each group in conf order (:519-576);
//   * field f starts at tuplet*ceil(end/tuplet) with tuplet = entry size times
//     dims[0] for multi-dimensional fields and the entry size otherwise,
//     computed in size_t so that the first field (end = 0) lands at
//     (SIZE_MAX/tuplet + 1)*tuplet mod 2^64 (:373-382).
// Where the reference has undefined behaviour the builder reports an error
// instead (missing smoothingradius / gridbounds field, zero tuplet size), or
// takes the benign reading (fewer defaults than the field holds: the rest stay
// zero; a dimension naming a field without a value: 0).
#include <cinttypes>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "host_common.h"

namespace sphhip {
namespace {

thread_local int t_build_status = 0;

std::mutex g_kind_mutex;
std::map<const void *, std::vector<FieldKind>> g_kinds;

struct FieldSpec {
    std::string name;
    std::vector<unsigned int> dims;
    std::string type;
    std::vector<unsigned char> value;   // parsed defaults; empty = none
    bool has_value = false;
    bool last_dim_is_pnum = false;
    FieldKind kind = FieldKind::Other;
};

// Pieces of `s` between runs of delimiter characters (strtok semantics).
std::vector<std::string> tokens(const std::string &s, const char *delims) {
    std::vector<std::string> out;
    size_t i = 0;
    const size_t n = s.size();
    while (i < n) {
        while (i < n && std::strchr(delims, s[i]) != nullptr) ++i;
        if (i >= n) break;
        size_t j = i;
        while (j < n && std::strchr(delims, s[j]) == nullptr) ++j;
        out.push_back(s.substr(i, j - i));
        i = j;
    }
    return out;
}

size_t type_size(const std::string &t) {       // src/stringly.c:14-20,45-61
    if (t == "char") return 1;
    if (t == "unsigned int" || t == "uint") return 4;
    if (t == "int") return 4;
    if (t == "float" || t == "float32") return 4;
    if (t == "double" || t == "float64") return 8;
    return 0;
}

FieldType type_code(const std::string &t) {    // src/stringly.c:28-43
    if (t == "unsigned int" || t == "uint") return T_UINT;
    if (t == "int") return T_INT;
    if (t == "float" || t == "float32") return T_FLOAT32;
    if (t == "double" || t == "float64") return T_FLOAT64;
    return T_CHAR;
}

const FieldSpec *find_spec(const std::vector<FieldSpec> &list, const std::string &name) {
    for (const FieldSpec &f : list)
        if (f.name == name) return &f;
    return nullptr;
}

// First 4 bytes of a field's default as unsigned int; UINT_MAX when absent.
unsigned int spec_value_u32(const std::vector<FieldSpec> &list, const std::string &name,
                            bool *missing_value) {
    const FieldSpec *f = find_spec(list, name);
    if (f == nullptr) return UINT_MAX;
    if (!f->has_value || f->value.size() < 4) {
        if (missing_value) *missing_value = true;
        return 0;
    }
    unsigned int v;
    std::memcpy(&v, f->value.data(), 4);
    return v;
}

template <typename T>
void append_value(std::vector<unsigned char> &bytes, T v) {
    const unsigned char *p = reinterpret_cast<const unsigned char *>(&v);
    bytes.insert(bytes.end(), p, p + sizeof(T));
}

bool parse_line(const std::string &line, const std::vector<FieldSpec> &list, FieldSpec *out) {
    constexpr size_t kMaxDims = 5, kMaxType = 64, kMaxValueBytes = 256;
    FieldSpec spec;
    bool named = false;
    std::string type;
    std::vector<unsigned char> value;
    std::vector<std::string> segments = tokens(line, "|");
    for (size_t seg = 0; seg < segments.size(); ++seg) {
        std::vector<std::string> words = tokens(segments[seg], ", ");
        for (size_t w = 0; w < words.size(); ++w) {
            const std::string &word = words[w];
            if (seg == 0 && w == 0) {
                spec.name = word;
                named = true;
            } else if (seg == 1) {
                if (spec.dims.size() == kMaxDims) {
                    note(1, "%s: Over maximum number of dimensions\n", spec.name.c_str());
                    break;
                }
                unsigned int dim = static_cast<unsigned int>(std::strtol(word.c_str(), nullptr, 0));
                bool named_dim = false;
                if (dim == 0 && !list.empty()) {
                    dim = spec_value_u32(list, word, nullptr);
                    if (dim == UINT_MAX) dim = 0;
                    named_dim = true;
                }
                spec.dims.push_back(dim);
                spec.last_dim_is_pnum = named_dim && word == "pnum";
            } else if (seg == 2) {
                if (type.size() + word.size() + 1 >= kMaxType) {
                    note(1, "%s: Over maximum type string length\n", spec.name.c_str());
                    break;
                }
                if (!type.empty()) type += ' ';
                type += word;
            } else if (seg == 3) {
                size_t need = 0;
                if (type == "double") need = sizeof(double);
                else if (type == "float") need = sizeof(float);
                else if (type == "unsigned int") need = sizeof(unsigned int);
                else if (type == "int") need = sizeof(int);
                if (need == 0) continue;
                if (value.size() + need > kMaxValueBytes) {
                    note(1, "%s: Over maximum default value length\n", spec.name.c_str());
                    break;
                }
                if (type == "double") append_value(value, std::strtod(word.c_str(), nullptr));
                else if (type == "float") append_value(value, std::strtof(word.c_str(), nullptr));
                else if (type == "unsigned int")
                    append_value(value, static_cast<unsigned int>(std::strtol(word.c_str(), nullptr, 0)));
                else append_value(value, static_cast<int>(std::strtol(word.c_str(), nullptr, 0)));
            }
        }
    }
    if (!named) return false;
    spec.type = type;
    spec.has_value = !value.empty();
    spec.value = std::move(value);
    spec.kind = spec.last_dim_is_pnum ? FieldKind::Particle : FieldKind::Other;
    *out = std::move(spec);
    return true;
}

// Appends the five grid arrays.  Returns false on a hard error.
bool add_grid_arrays(std::vector<FieldSpec> &list) {
    const FieldSpec *h = find_spec(list, "smoothingradius");
    const FieldSpec *gb = find_spec(list, "gridbounds");
    if (h == nullptr || gb == nullptr) {
        note(2, "psdata specification needs smoothingradius and gridbounds fields\n");
        return false;
    }
    if (!h->has_value) { note(1, "No value for smoothingradius\n"); return true; }
    if (h->type != "double" && h->type != "float") {
        note(2, "smoothingradius must be float or double\n");
        return false;
    }
    if (!gb->has_value) { note(1, "No value for gridbounds\n"); return true; }
    if (gb->type != "double" && gb->type != "float") {
        note(2, "gridbounds must be float or double\n");
        return false;
    }
    if ((h->type == "double") != (gb->type == "double")) {
        note(2, "smoothingradius and gridbounds must have the same type\n");
        return false;
    }
    bool missing = false;
    unsigned int pnum = spec_value_u32(list, "pnum", &missing);
    if (missing) { note(2, "pnum has no value\n"); return false; }
    if (pnum == UINT_MAX) { note(1, "No value for pnum\n"); return true; }

    double extent[3];
    double radius;
    if (h->type == "double") {
        double b[6] = {0, 0, 0, 0, 0, 0};
        std::memcpy(b, gb->value.data(), std::min(gb->value.size(), sizeof b));
        std::memcpy(&radius, h->value.data(), sizeof(double));
        for (int a = 0; a < 3; ++a) extent[a] = std::fabs(b[2 * a + 1] - b[2 * a]);
    } else {
        float b[6] = {0, 0, 0, 0, 0, 0};
        std::memcpy(b, gb->value.data(), std::min(gb->value.size(), sizeof b));
        float rf;
        std::memcpy(&rf, h->value.data(), sizeof(float));
        radius = static_cast<double>(rf);
        for (int a = 0; a < 3; ++a)
            extent[a] = std::fabs(static_cast<double>(b[2 * a + 1] - b[2 * a]));
    }
    unsigned int res[3];
    for (int a = 0; a < 3; ++a) {
        double r = extent[a] / radius;
        if (!(r >= 0.0) || r >= 4294967296.0) {
            note(2, "grid resolution along axis %d is not representable\n", a);
            return false;
        }
        res[a] = static_cast<unsigned int>(r);
    }

    FieldSpec f;
    f.type = "unsigned int";

    f.name = "gridres"; f.dims = {3}; f.kind = FieldKind::Other;
    f.value.clear();
    for (unsigned int r : res) append_value(f.value, r);
    f.has_value = true;
    list.push_back(f);

    f.value.clear(); f.has_value = false;
    f.name = "gridcell"; f.dims = {pnum}; f.kind = FieldKind::Particle;
    list.push_back(f);
    f.name = "gridcount"; f.dims = {res[0], res[1], res[2]}; f.kind = FieldKind::Cell;
    list.push_back(f);
    f.name = "celloffset"; f.kind = FieldKind::Cell;
    list.push_back(f);
    f.name = "cellparticles"; f.dims = {pnum}; f.kind = FieldKind::Slot;
    list.push_back(f);
    return true;
}

// Regroup as double, float, unsigned int, then everything else.
std::vector<FieldSpec> group_by_type(std::vector<FieldSpec> list) {
    static const char *const order[] = {"double", "float", "unsigned int"};
    std::vector<FieldSpec> out;
    std::vector<bool> taken(list.size(), false);
    for (const char *t : order)
        for (size_t i = 0; i < list.size(); ++i)
            if (!taken[i] && list[i].type == t) { out.push_back(list[i]); taken[i] = true; }
    for (size_t i = 0; i < list.size(); ++i)
        if (!taken[i]) out.push_back(list[i]);
    return out;
}

template <typename T>
T *alloc_array(size_t count) {
    return static_cast<T *>(std::malloc(count == 0 ? 1 : count * sizeof(T)));
}

bool populate(psdata *data, const std::vector<FieldSpec> &list) {
    const unsigned int nf = static_cast<unsigned int>(list.size());
    std::vector<unsigned int> entry(nf), sizes(nf), offsets(nf), name_off(nf), dim_off(nf);
    size_t names_len = 0, dims_len = 0, end = 0;
    for (unsigned int f = 0; f < nf; ++f) {
        const FieldSpec &s = list[f];
        name_off[f] = static_cast<unsigned int>(names_len);
        names_len += s.name.size() + 1;
        dim_off[f] = static_cast<unsigned int>(dims_len);
        dims_len += s.dims.size();
        entry[f] = static_cast<unsigned int>(type_size(s.type));
        unsigned int bytes = entry[f];
        for (unsigned int d : s.dims) bytes *= d;
        sizes[f] = bytes;
        size_t tuplet = s.dims.size() > 1 ? static_cast<size_t>(entry[f]) * s.dims[0] : entry[f];
        if (tuplet == 0) {
            note(2, "%s: zero-sized tuple (type '%s')\n", s.name.c_str(), s.type.c_str());
            return false;
        }
        size_t off = tuplet * ((end - 1) / tuplet + 1);   // wraps to 0 (or 8, ...) for end == 0
        offsets[f] = static_cast<unsigned int>(off);
        end = static_cast<size_t>(offsets[f]) + bytes;
    }

    data->num_fields = nf;
    char *names = alloc_array<char>(names_len);
    data->names_offsets = alloc_array<unsigned int>(nf);
    data->dimensions = alloc_array<unsigned int>(dims_len);
    data->num_dimensions = alloc_array<unsigned int>(nf);
    data->dimensions_offsets = alloc_array<unsigned int>(nf);
    data->entry_sizes = alloc_array<unsigned int>(nf);
    data->data_sizes = alloc_array<unsigned int>(nf);
    data->data_offsets = alloc_array<unsigned int>(nf);
    data->data_ptrs = alloc_array<DataPtr>(nf);
    data->data = std::calloc(end == 0 ? 1 : end, 1);

    std::vector<FieldKind> kinds(nf);
    for (unsigned int f = 0; f < nf; ++f) {
        const FieldSpec &s = list[f];
        std::memcpy(names + name_off[f], s.name.c_str(), s.name.size() + 1);
        data->names_offsets[f] = name_off[f];
        data->num_dimensions[f] = static_cast<unsigned int>(s.dims.size());
        data->dimensions_offsets[f] = dim_off[f];
        for (size_t d = 0; d < s.dims.size(); ++d) data->dimensions[dim_off[f] + d] = s.dims[d];
        data->entry_sizes[f] = entry[f];
        data->data_sizes[f] = sizes[f];
        data->data_offsets[f] = offsets[f];
        char *field = static_cast<char *>(data->data) + offsets[f];
        if (s.has_value) std::memcpy(field, s.value.data(), std::min<size_t>(s.value.size(), sizes[f]));
        data->data_ptrs[f].type = type_code(s.type);
        data->data_ptrs[f].ptr = field;
        kinds[f] = s.kind;
    }
    data->names = names;
    data->num_host_fields = 0;
    data->host_names = nullptr;
    data->host_data = nullptr;
    data->host_data_size = nullptr;

    std::lock_guard<std::mutex> lock(g_kind_mutex);
    g_kinds[data->data] = std::move(kinds);
    return true;
}

}  // namespace

std::vector<FieldKind> field_kinds_for(const void *blob) {
    std::lock_guard<std::mutex> lock(g_kind_mutex);
    auto it = g_kinds.find(blob);
    return it == g_kinds.end() ? std::vector<FieldKind>() : it->second;
}

void set_build_status(int status) { t_build_status = status; }

}  // namespace sphhip

using sphhip::FieldSpec;

extern "C" int psdata_last_build_status(void) { return sphhip::t_build_status; }

extern "C" void build_psdata_from_string(psdata *data, const char *string) {
    using namespace sphhip;
    std::memset(data, 0, sizeof *data);
    t_build_status = 0;
    if (string == nullptr) { t_build_status = 1; note(2, "no psdata specification\n"); return; }
    std::vector<FieldSpec> list;
    for (const std::string &line : tokens(string, "\n")) {
        if (!line.empty() && line[0] == '#') continue;
        FieldSpec spec;
        if (parse_line(line, list, &spec)) list.push_back(std::move(spec));
    }
    if (list.empty()) { t_build_status = 1; note(2, "empty psdata specification\n"); return; }
    if (!add_grid_arrays(list)) { t_build_status = 1; return; }
    if (!populate(data, group_by_type(std::move(list)))) {
        std::memset(data, 0, sizeof *data);
        t_build_status = 1;
    }
}

extern "C" void build_psdata(psdata *data, const char *path) {
    std::FILE *f = std::fopen(path, "rb");
    if (!f) {
        note(1, "Could not read file %s\n", path);
        std::memset(data, 0, sizeof *data);
        sphhip::t_build_status = 1;
        return;
    }
    std::fseek(f, 0, SEEK_END);
    long len = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::string text(len > 0 ? static_cast<size_t>(len) : 0, '\0');
    size_t got = std::fread(&text[0], 1, text.size(), f);
    std::fclose(f);
    text.resize(got);
    if (!text.empty()) text.pop_back();   // the reference overwrites the last byte with NUL
    build_psdata_from_string(data, text.c_str());
}

extern "C" int get_field_psdata(psdata data, const char *name) {
    for (unsigned int i = 0; i < data.num_fields; ++i)
        if (std::strcmp(data.names + data.names_offsets[i], name) == 0) return static_cast<int>(i);
    return -1;
}

extern "C" void set_field_psdata(psdata *data, const char *name, void *field, unsigned int size,
                                 unsigned int offset) {
    int i = get_field_psdata(*data, name);
    if (i != -1) std::memcpy(static_cast<char *>(data->data) + data->data_offsets[i] + offset, field, size);
}

extern "C" unsigned int psdata_names_size(psdata data) {
    if (data.num_fields == 0) return 0;
    unsigned int last = data.num_fields - 1;
    return data.names_offsets[last] +
           static_cast<unsigned int>(std::strlen(data.names + data.names_offsets[last])) + 1;
}

extern "C" unsigned int psdata_dimensions_size(psdata data) {
    if (data.num_fields == 0) return 0;
    unsigned int last = data.num_fields - 1;
    return (data.dimensions_offsets[last] + data.num_dimensions[last]) * sizeof(unsigned int);
}

extern "C" unsigned int psdata_data_size(psdata data) {
    if (data.num_fields == 0) return 0;
    unsigned int last = data.num_fields - 1;
    return data.data_offsets[last] + data.data_sizes[last];
}

extern "C" int create_host_field_psdata(psdata *data, const char *name, void *field, unsigned int size) {
    unsigned int k = data->num_host_fields;
    char **names = static_cast<char **>(std::malloc((k + 1) * sizeof(char *)));
    void **ptrs = static_cast<void **>(std::malloc((k + 1) * sizeof(void *)));
    unsigned int *sizes = static_cast<unsigned int *>(std::malloc((k + 1) * sizeof(unsigned int)));
    for (unsigned int i = 0; i < k; ++i) {
        names[i] = data->host_names[i];
        ptrs[i] = data->host_data[i];
        sizes[i] = data->host_data_size[i];
    }
    if (k > 0) {
        std::free(data->host_names);
        std::free(data->host_data);
        std::free(data->host_data_size);
    }
    names[k] = static_cast<char *>(std::malloc(std::strlen(name) + 1));
    std::strcpy(names[k], name);
    ptrs[k] = field;
    sizes[k] = size;
    data->num_host_fields = k + 1;
    data->host_names = names;
    data->host_data = ptrs;
    data->host_data_size = sizes;
    return static_cast<int>(k);
}

extern "C" int get_host_field_psdata(psdata *data, const char *name) {
    for (unsigned int i = 0; i < data->num_host_fields; ++i)
        if (std::strcmp(data->host_names[i], name) == 0) return static_cast<int>(i);
    return -1;
}

extern "C" void display_psdata(psdata data, const char *const *mask) {
    if (mask != nullptr) return;   // the reference only implements the unmasked dump
    for (unsigned int f = 0; f < data.num_fields; ++f) {
        note(2, "%s:\n\n", data.names + data.names_offsets[f]);
        const unsigned int *dims = data.dimensions + data.dimensions_offsets[f];
        unsigned int nd = data.num_dimensions[f], es = data.entry_sizes[f];
        size_t count = 1;
        for (unsigned int d = 0; d < nd && d < 3; ++d) count *= dims[d];
        const char *base = static_cast<const char *>(data.data) + data.data_offsets[f];
        for (size_t e = 0; e < count; ++e) {
            if (es == 8) {
                double v;
                std::memcpy(&v, base + e * 8, 8);
                note(2, "%g, ", v);
            } else if (es == 4) {
                unsigned int v;
                std::memcpy(&v, base + e * 4, 4);
                note(2, "%u, ", v);
            }
            if (nd >= 2 && (e + 1) % dims[0] == 0) note(2, "\n");
        }
        note(2, "\n\n");
    }
}

extern "C" void free_psdata(psdata *data) {
    {
        std::lock_guard<std::mutex> lock(sphhip::g_kind_mutex);
        sphhip::g_kinds.erase(data->data);
    }
    std::free(const_cast<char *>(data->names));
    std::free(data->names_offsets);
    std::free(data->dimensions);
    std::free(data->num_dimensions);
    std::free(data->dimensions_offsets);
    std::free(data->entry_sizes);
    std::free(data->data);
    std::free(data->data_sizes);
    std::free(data->data_offsets);
    std::free(data->data_ptrs);
    for (unsigned int i = 0; i < data->num_host_fields; ++i) {
        std::free(data->host_data[i]);
        std::free(data->host_names[i]);
    }
    if (data->num_host_fields > 0) {
        std::free(data->host_names);
        std::free(data->host_data);
        std::free(data->host_data_size);
    }
    note(1, "freed host data\n");
    std::memset(data, 0, sizeof *data);
}
