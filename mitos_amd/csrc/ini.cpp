// .ini reader for the keys of the IC path, with the value rules of the
// reference's minIni getters (parser/minIni.c:335-403): sections and keys are
// case-insensitive, ';' and '#' start a comment outside double quotes,
// surrounding quotes are removed, integers are decimal unless the second
// character is 'x' (hexadecimal), reals go through strtod.
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include <string>

#include "../../include/mitos_hip.h"

namespace {

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && isspace((unsigned char)s[a])) a++;
    while (b > a && isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

// value of [section] key, or "" when absent
std::string ini_value(FILE* f, const char* section, const char* key) {
    rewind(f);
    char line[1024];
    bool in = false;
    while (fgets(line, sizeof(line), f)) {
        std::string s = trim(line);
        if (s.empty()) continue;
        if (s[0] == '[') {
            size_t e = s.find(']');
            std::string name = trim(s.substr(1, e == std::string::npos ? std::string::npos : e - 1));
            in = strcasecmp(name.c_str(), section) == 0;
            continue;
        }
        if (!in || s[0] == ';' || s[0] == '#') continue;
        size_t eq = s.find_first_of("=:");
        if (eq == std::string::npos) continue;
        if (strcasecmp(trim(s.substr(0, eq)).c_str(), key) != 0) continue;
        std::string v = s.substr(eq + 1);
        // strip a trailing comment outside quotes (minIni.c:200-213)
        bool str = false;
        size_t i = 0;
        for (; i < v.size(); i++) {
            if (v[i] == '"') {
                if (i + 1 < v.size() && v[i + 1] == '"') i++;
                else str = !str;
            } else if (v[i] == '\\' && i + 1 < v.size() && v[i + 1] == '"') {
                i++;
            } else if ((v[i] == ';' || v[i] == '#') && !str) {
                break;
            }
        }
        v = trim(v.substr(0, i));
        if (v.size() >= 2 && v.front() == '"' && v.back() == '"') v = v.substr(1, v.size() - 2);
        return v;
    }
    return "";
}

long getl(FILE* f, const char* sec, const char* key, long def) {
    std::string v = ini_value(f, sec, key);
    if (v.empty()) return def;
    const int base = (v.size() >= 2 && toupper((unsigned char)v[1]) == 'X') ? 16 : 10;
    return strtol(v.c_str(), nullptr, base);
}

double getd(FILE* f, const char* sec, const char* key, double def) {
    std::string v = ini_value(f, sec, key);
    return v.empty() ? def : strtod(v.c_str(), nullptr);
}

}  // namespace

extern "C" int mitos_read_ini(const char* fname, struct mitos_ic_config* cfg) {
    if (!fname || !cfg) return MITOS_ERR_ARG;
    FILE* f = fopen(fname, "r");
    if (!f) return MITOS_ERR_ARG;
    cfg->seed = (uint64_t)getl(f, "Random", "Seed", 1);           // src/input.c:29
    cfg->N = (int)getl(f, "Box", "GridSize", 64);                   // src/input.c:31
    cfg->boxlen = getd(f, "Box", "BoxLen", 1.0);                    // src/input.c:33
    cfg->n_s = getd(f, "Cosmology", "n_s", 0.965);
    cfg->sigma8 = getd(f, "Spectrum", "Sigma8", 0.81);
    cfg->gamma = getd(f, "Spectrum", "Gamma", 0.21);
    cfg->bins = (int)getl(f, "Read", "PowerSpectrumBins", 50);     // src/input.c:46
    cfg->run2lpt = (int)getl(f, "ParticleType_0", "Run2LPT", 0);
    cfg->f1 = 1.0;                                                  // src/mitos.c:385
    cfg->f2 = -3.0 / 7.0;
    const long M = getl(f, "ParticleType_0", "CubeRootNumber", cfg->N);
    std::string prec = ini_value(f, "GPU", "Precision");
    fclose(f);
    cfg->precision = (strcasecmp(prec.c_str(), "FP32") == 0) ? MITOS_FP32 : MITOS_FP64;
    if (M < 1 || M > 2048) return MITOS_ERR_ARG;
    cfg->M = M == cfg->N ? 0 : (int)M;  // src/particle.c:87-129 lattice of M^3 particles
    return MITOS_OK;
}
