// mc_io.cpp -- parameter output (CSV) and binary checkpoints of the EM engine.
//
// The CSV files are those of Communicate::output_parameters
// (parallel_em/communicator.cpp:415-516) for a single worker; checkpoints
// hold every table's numerators and denominators so that training can resume
// from them.

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mclicks.h"
#include "mc_internal.h"

namespace mc {
namespace {

constexpr float kParamDefault = 1.f / 2.f;   // PARAM_DEF_NUM / PARAM_DEF_DENOM, utils/definitions.h:42-43
constexpr char kCheckpointMagic[8] = {'M', 'C', 'L', 'K', 'P', 'R', 'M', '1'};

struct TableDesc {
    std::string name;
    bool is_public;
    int64_t size;
};

// Tables of a model type in mc_get_table_info order.
std::vector<TableDesc> describe(int type, int64_t n_qd) {
    switch (type) {
        case MC_PBM: return {{"attractiveness", false, n_qd}, {"examination", true, kMaxSerp}};
        case MC_UBM: return {{"attractiveness", false, n_qd}, {"examination", true, 55}};
        case MC_CCM: return {{"attractiveness", false, n_qd}, {"continuation", true, 3}};
        case MC_DBN:
            return {{"attractiveness", false, n_qd}, {"satisfaction", false, n_qd}, {"continuation", true, 1}};
    }
    return {};
}

// Param::value (click_models/param.cu:75-87).
inline float param_value(float num, float den) {
    const float v = num / den;
    return v < 1.f - 0.000001f ? v : 1.f - 0.000001f;
}

// std::ostream << float with the default format: %g, 6 significant digits.
void append_ostream_float(std::string& out, float v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%g", (double)v);
    out += buf;
}

// std::to_string(float): %f of the value promoted to double.
void append_to_string_float(std::string& out, float v) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "%f", (double)v);
    out += buf;
}

bool write_file(const std::string& path, const std::string& bytes) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
    return std::fclose(f) == 0 && ok;
}

}  // namespace
}  // namespace mc

using namespace mc;

extern "C" {

mc_status mc_write_parameters_csv(const mc_dataset* ds, const char* prefix, int model_type,
                                  const float* const* num, const float* const* den) {
    if (!ds || !prefix || !num || !den || model_type < 0 || model_type > 3) {
        set_error("invalid argument");
        return MC_ERR_INVALID_ARG;
    }
    const HostDataset& d = ds->d;
    const std::vector<TableDesc> tabs = describe(model_type, d.n_qd);
    for (size_t t = 0; t < tabs.size(); t++) {
        if (!num[t] || !den[t]) {
            set_error("null table " + std::to_string(t));
            return MC_ERR_INVALID_ARG;
        }
    }
    const std::string base(prefix);
    for (size_t t = 0; t < tabs.size(); t++) {
        const TableDesc& td = tabs[t];
        const std::string path = base + "_" + td.name + ".csv";
        std::string bytes;
        if (td.is_public) {
            // communicator.cpp:437-445
            bytes = "rank, " + td.name + "\n";
            for (int64_t r = 0; r < td.size; r++) {
                bytes += std::to_string(r);
                bytes += ", ";
                append_ostream_float(bytes, param_value(num[t][r], den[t][r]));
                bytes += "\n";
            }
        } else {
            // communicator.cpp:448-500: the header goes in first, then the
            // data is written at offset 0 over it.
            const std::string header = "query, document, " + td.name + "\n";
            const int64_t n = (int64_t)d.train_query.size();
            bytes.reserve((size_t)n * kMaxSerp * 32);
            for (int64_t s = 0; s < n; s++) {
                const std::string q = std::to_string(d.train_query[s]);
                for (int r = 0; r < kMaxSerp; r++) {
                    const int32_t p = d.train_pidx[s * kMaxSerp + r];
                    const float v = p != -1 ? param_value(num[t][p], den[t][p]) : kParamDefault;
                    bytes += q;
                    bytes += ", ";
                    bytes += std::to_string(d.train_docs[s * kMaxSerp + r]);
                    bytes += ", ";
                    append_to_string_float(bytes, v);
                    bytes += "\n";
                }
            }
            if (bytes.size() < header.size()) bytes += header.substr(bytes.size());
        }
        if (!write_file(path, bytes)) {
            set_error("cannot write " + path);
            return MC_ERR_IO;
        }
    }
    return MC_OK;
}

mc_status mc_output_parameters(mc_model* m, const mc_dataset* ds, const char* prefix) {
    if (!m || !ds || !prefix) {
        set_error("invalid argument");
        return MC_ERR_INVALID_ARG;
    }
    int n_tables = 0;
    mc_status st = mc_get_n_tables(m, &n_tables);
    if (st != MC_OK) return st;
    int type = -1;
    if ((st = mc_get_model_type(m, &type)) != MC_OK) return st;
    std::vector<std::vector<float>> num((size_t)n_tables), den((size_t)n_tables);
    std::vector<const float*> pn((size_t)n_tables), pd((size_t)n_tables);
    for (int t = 0; t < n_tables; t++) {
        int64_t size = 0;
        int is_public = 0;
        const char* name = nullptr;
        if ((st = mc_get_table_info(m, t, &name, &is_public, &size)) != MC_OK) return st;
        if (!is_public && size != ds->d.n_qd) {
            set_error("the dataset's n_qd does not match the model's tables");
            return MC_ERR_INVALID_ARG;
        }
        num[t].resize((size_t)size);
        den[t].resize((size_t)size);
        if ((st = mc_get_parameters(m, t, num[t].data(), den[t].data())) != MC_OK) return st;
        pn[t] = num[t].data();
        pd[t] = den[t].data();
    }
    return mc_write_parameters_csv(ds, prefix, type, pn.data(), pd.data());
}

mc_status mc_save_parameters(mc_model* m, const char* path) {
    if (!m || !path) {
        set_error("invalid argument");
        return MC_ERR_INVALID_ARG;
    }
    int n_tables = 0;
    mc_status st = mc_get_n_tables(m, &n_tables);
    if (st != MC_OK) return st;
    int type = -1;
    if ((st = mc_get_model_type(m, &type)) != MC_OK) return st;
    std::string bytes(kCheckpointMagic, sizeof kCheckpointMagic);
    auto put = [&](const void* p, size_t n) { bytes.append((const char*)p, n); };
    const int32_t ty = type, nt = n_tables;
    put(&ty, 4);
    put(&nt, 4);
    for (int t = 0; t < n_tables; t++) {
        int64_t size = 0;
        int is_public = 0;
        const char* name = nullptr;
        if ((st = mc_get_table_info(m, t, &name, &is_public, &size)) != MC_OK) return st;
        std::vector<float> num((size_t)size), den((size_t)size);
        if ((st = mc_get_parameters(m, t, num.data(), den.data())) != MC_OK) return st;
        char nm[32] = {0};
        std::strncpy(nm, name, sizeof nm - 1);
        put(nm, sizeof nm);
        put(&size, 8);
        put(num.data(), (size_t)size * 4);
        put(den.data(), (size_t)size * 4);
    }
    if (!write_file(path, bytes)) {
        set_error(std::string("cannot write ") + path);
        return MC_ERR_IO;
    }
    return MC_OK;
}

mc_status mc_load_parameters(mc_model* m, const char* path) {
    if (!m || !path) {
        set_error("invalid argument");
        return MC_ERR_INVALID_ARG;
    }
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        set_error(std::string("cannot open ") + path);
        return MC_ERR_IO;
    }
    auto fail = [&](const std::string& msg, mc_status code) {
        std::fclose(f);
        set_error(msg);
        return code;
    };
    char magic[8];
    int32_t nt = 0;
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kCheckpointMagic, 8) != 0)
        return fail(std::string(path) + " is not a parameter checkpoint", MC_ERR_INVALID_ARG);
    int n_tables = 0, type = -1;
    mc_status st = mc_get_n_tables(m, &n_tables);
    if (st == MC_OK) st = mc_get_model_type(m, &type);
    if (st != MC_OK) return fail(mc_last_error(), st);
    int32_t ty = -1;
    if (std::fread(&ty, 4, 1, f) != 1 || ty != type)
        return fail("checkpoint model type does not match the model", MC_ERR_INVALID_ARG);
    if (std::fread(&nt, 4, 1, f) != 1 || nt != n_tables)
        return fail("checkpoint table count does not match the model", MC_ERR_INVALID_ARG);
    std::vector<std::vector<float>> num((size_t)nt), den((size_t)nt);
    for (int t = 0; t < nt; t++) {
        char nm[32];
        int64_t size = 0, want = 0;
        int is_public = 0;
        const char* name = nullptr;
        if ((st = mc_get_table_info(m, t, &name, &is_public, &want)) != MC_OK) return fail(mc_last_error(), st);
        if (std::fread(nm, 1, 32, f) != 32 || std::fread(&size, 8, 1, f) != 1)
            return fail("truncated checkpoint", MC_ERR_INVALID_ARG);
        nm[31] = 0;
        if (std::strcmp(nm, name) != 0 || size != want)
            return fail("checkpoint table " + std::to_string(t) + " (" + nm + ", " + std::to_string(size) +
                            ") does not match the model's (" + name + ", " + std::to_string(want) + ")",
                        MC_ERR_INVALID_ARG);
        num[t].resize((size_t)size);
        den[t].resize((size_t)size);
        if (std::fread(num[t].data(), 4, (size_t)size, f) != (size_t)size ||
            std::fread(den[t].data(), 4, (size_t)size, f) != (size_t)size)
            return fail("truncated checkpoint", MC_ERR_INVALID_ARG);
    }
    std::fclose(f);
    for (int t = 0; t < nt; t++)
        if ((st = mc_set_parameters(m, t, num[t].data(), den[t].data())) != MC_OK) return st;
    return MC_OK;
}

}  // extern "C"
