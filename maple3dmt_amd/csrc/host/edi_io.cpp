// EDI reader / writer; see edi_io.h for the conventions and the reference
// functions it mirrors (src/io/edi_io.cpp:156-619).
#include "edi_io.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <map>
#include <stdexcept>

namespace m3d {
namespace io {

namespace {

constexpr double kZUnit = 4.0 * M_PI * 1e-4;  // mV/km/nT -> ohm (mu0 * 1e3)

std::string strip(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

std::string upper(std::string s) {
    for (auto& ch : s) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
    return s;
}

// Value of KEY=value (or KEY =value) on a header line; quoted values keep
// their spaces.  Empty when the key is absent.
std::string value_of(const std::string& line, const std::string& key) {
    size_t p = 0;
    while ((p = line.find(key, p)) != std::string::npos) {
        const bool start_ok = p == 0 || !std::isalnum(static_cast<unsigned char>(line[p - 1]));
        size_t q = p + key.size();
        while (q < line.size() && (line[q] == ' ' || line[q] == '\t')) ++q;
        if (start_ok && q < line.size() && line[q] == '=') {
            ++q;
            while (q < line.size() && (line[q] == ' ' || line[q] == '\t')) ++q;
            if (q < line.size() && line[q] == '"') {
                const size_t e = line.find('"', q + 1);
                if (e != std::string::npos) return line.substr(q + 1, e - q - 1);
            }
            const size_t e = line.find_first_of(" \t\r\n", q);
            return line.substr(q, (e == std::string::npos ? line.size() : e) - q);
        }
        p += key.size();
    }
    return "";
}

// Degrees from "+35:32:28.7" (sign applies to the whole angle) or "35.541".
double coordinate(const std::string& s) {
    if (s.empty()) return 0.0;
    if (s.find(':') == std::string::npos) return std::strtod(s.c_str(), nullptr);
    double sign = 1.0;
    size_t i = 0;
    if (s[0] == '+' || s[0] == '-') {
        sign = s[0] == '-' ? -1.0 : 1.0;
        i = 1;
    }
    double part[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 3 && i <= s.size(); ++k) {
        const size_t e = s.find(':', i);
        part[k] = std::strtod(s.substr(i, e == std::string::npos ? std::string::npos : e - i).c_str(), nullptr);
        if (e == std::string::npos) break;
        i = e + 1;
    }
    return sign * (part[0] + part[1] / 60.0 + part[2] / 3600.0);
}

// The tag of a section line ">TAG rest": the first token after '>', cut at
// whitespace or '/'.
std::string tag_of(const std::string& line) {
    size_t e = 1;
    while (e < line.size() && line[e] != ' ' && line[e] != '\t' && line[e] != '/') ++e;
    return upper(line.substr(1, e - 1));
}

// Count on a >FREQ line: NFREQ=n or "//n"; -1 when neither is there.
int count_of(const std::string& line) {
    const std::string v = value_of(upper(line), "NFREQ");
    if (!v.empty()) return std::atoi(v.c_str());
    const size_t p = line.find("//");
    if (p != std::string::npos) {
        const std::string r = strip(line.substr(p + 2));
        if (!r.empty() && std::isdigit(static_cast<unsigned char>(r[0]))) return std::atoi(r.c_str());
    }
    return -1;
}

// Up to n numbers from the lines after `pos`, stopping before the next
// section ('>') or comment (';') line; `pos` ends on the last line read.
std::vector<double> numbers(const std::vector<std::string>& lines, size_t& pos, int n) {
    std::vector<double> out;
    while (static_cast<int>(out.size()) < n && pos + 1 < lines.size()) {
        const std::string& l = lines[pos + 1];
        if (!l.empty() && (l[0] == '>' || l[0] == ';')) break;
        ++pos;
        const char* p = l.c_str();
        char* end = nullptr;
        while (static_cast<int>(out.size()) < n) {
            const double v = std::strtod(p, &end);
            if (end == p) break;
            out.push_back(v);
            p = end;
        }
    }
    return out;
}

// Tipper tags the reference accepts (TXR, TXR.EXP, TXYR.EXP, TX.VAR, TXVAR ...).
enum Slot { kNone = -1, kRe = 0, kIm = 1, kVar = 2 };

}  // namespace

EDIRecord read_edi(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open EDI file " + path);
    std::vector<std::string> lines;
    for (std::string l; std::getline(in, l);) lines.push_back(strip(l));

    EDIRecord rec;
    int nfreq = -1;
    // comp[c][slot] for c = ZXX, ZXY, ZYX, ZYY, TX, TY
    std::vector<double> comp[6][3];
    bool present[6] = {false, false, false, false, false, false};
    const char* znames[4] = {"ZXX", "ZXY", "ZYX", "ZYY"};

    for (size_t i = 0; i < lines.size(); ++i) {
        const std::string& l = lines[i];
        if (l.empty() || l[0] != '>') continue;
        const std::string tag = tag_of(l);
        if (tag == "END") break;
        if (tag == "HEAD" || tag == "=MTSECT" || tag == "MTSECT") {
            // key=value lines up to the next section
            for (; i + 1 < lines.size() && (lines[i + 1].empty() || lines[i + 1][0] != '>'); ++i) {
                const std::string& h = lines[i + 1];
                if (h.empty() || h[0] == ';') continue;
                const std::string u = upper(h);
                if (tag == "HEAD") {
                    const std::string id = value_of(h, "DATAID");
                    if (!id.empty()) rec.station.name = id;
                    const std::string lat = value_of(u, "LAT");
                    if (!lat.empty()) { rec.station.lat = coordinate(lat); rec.station.has_geo = true; }
                    std::string lon = value_of(u, "LONG");
                    if (lon.empty()) lon = value_of(u, "LON");
                    if (!lon.empty()) { rec.station.lon = coordinate(lon); rec.station.has_geo = true; }
                    const std::string el = value_of(u, "ELEV");
                    if (!el.empty()) rec.station.z = std::strtod(el.c_str(), nullptr);
                } else {
                    const std::string nf = value_of(u, "NFREQ");
                    if (!nf.empty()) nfreq = std::atoi(nf.c_str());
                }
            }
            continue;
        }
        if (tag == "FREQ") {
            const int n = count_of(l);
            if (n > 0) nfreq = n;
            if (nfreq <= 0) throw std::runtime_error("EDI: no frequency count in " + path);
            rec.frequencies = numbers(lines, i, nfreq);
            if (static_cast<int>(rec.frequencies.size()) != nfreq)
                throw std::runtime_error("EDI: " + std::to_string(nfreq) + " frequencies announced, " +
                                         std::to_string(rec.frequencies.size()) + " found in " + path);
            continue;
        }
        if (rec.frequencies.empty()) continue;  // data sections count only after >FREQ
        int c = -1;
        Slot slot = kNone;
        for (int k = 0; k < 4; ++k) {
            const std::string z = znames[k];
            if (tag == z + "R") { c = k; slot = kRe; }
            else if (tag == z + "I") { c = k; slot = kIm; }
            else if (tag == z + ".VAR") { c = k; slot = kVar; }
        }
        if (c < 0) {
            std::string t = tag;
            if (t.size() > 4 && t.compare(t.size() - 4, 4, ".EXP") == 0) t.resize(t.size() - 4);
            for (int k = 0; k < 2; ++k) {
                const std::string a = k == 0 ? "TX" : "TY";
                const std::string b = k == 0 ? "TXY" : "TYY";
                if (t == a + "R" || t == b + "R") { c = 4 + k; slot = kRe; }
                else if (t == a + "I" || t == b + "I") { c = 4 + k; slot = kIm; }
                else if (t == a + ".VAR" || t == b + ".VAR" || t == a + "VAR") { c = 4 + k; slot = kVar; }
            }
        }
        if (c < 0) continue;
        comp[c][slot] = numbers(lines, i, nfreq);
        if (slot == kRe) present[c] = true;
    }
    if (rec.frequencies.empty()) throw std::runtime_error("EDI: no >FREQ section in " + path);

    const int nf = static_cast<int>(rec.frequencies.size());
    rec.responses.assign(nf, data::MTResponse{});
    for (int f = 0; f < nf; ++f) {
        data::MTResponse& r = rec.responses[f];
        data::Datum* d[6] = {&r.Zxx, &r.Zxy, &r.Zyx, &r.Zyy, &r.Tx, &r.Ty};
        for (int c = 0; c < 6; ++c) {
            if (!present[c]) {
                d[c]->weight = 0.0;
                continue;
            }
            auto at = [&](int slot) { return f < static_cast<int>(comp[c][slot].size()) ? comp[c][slot][f] : 0.0; };
            const double unit = c < 4 ? kZUnit : 1.0;
            d[c]->value = std::complex<double>(at(kRe) * unit, at(kIm) * unit);
            const double var = at(kVar);
            d[c]->error = var > 0.0 ? std::sqrt(var) * unit : (c < 4 ? kZUnit : 0.1);
            d[c]->weight = 1.0;
        }
    }
    return rec;
}

void write_edi(const std::string& path, const data::Station& st, const std::vector<double>& freqs,
               const std::vector<data::MTResponse>& resp) {
    if (resp.size() != freqs.size()) throw std::invalid_argument("EDI: one response per frequency");
    std::FILE* fp = std::fopen(path.c_str(), "w");
    if (!fp) throw std::runtime_error("cannot write EDI file " + path);
    const int nf = static_cast<int>(freqs.size());
    auto row = [&](auto value) {
        for (int f = 0; f < nf; ++f) {
            std::fprintf(fp, "  %.6e", value(f));
            if ((f + 1) % 6 == 0 || f == nf - 1) std::fprintf(fp, "\n");
        }
    };
    std::fprintf(fp, ">HEAD\n  DATAID=\"%s\"\n  LAT=%.6f\n  LONG=%.6f\n  ELEV=%.6f\n\n", st.name.c_str(), st.lat,
                 st.lon, st.z);
    std::fprintf(fp, ">FREQ NFREQ=%d\n", nf);
    row([&](int f) { return freqs[f]; });
    std::fprintf(fp, "\n");
    struct Sec { const char* name; data::Datum data::MTResponse::*m; double unit; };
    const Sec secs[6] = {{"ZXX", &data::MTResponse::Zxx, kZUnit}, {"ZXY", &data::MTResponse::Zxy, kZUnit},
                         {"ZYX", &data::MTResponse::Zyx, kZUnit}, {"ZYY", &data::MTResponse::Zyy, kZUnit},
                         {"TX", &data::MTResponse::Tx, 1.0},      {"TY", &data::MTResponse::Ty, 1.0}};
    for (const Sec& s : secs) {
        bool any = false;
        for (int f = 0; f < nf; ++f) any = any || (resp[f].*(s.m)).weight > 0.0;
        if (!any) continue;   // a component without data gets no section
        std::fprintf(fp, ">%sR ROT=0.0\n", s.name);
        row([&](int f) { return (resp[f].*(s.m)).value.real() / s.unit; });
        std::fprintf(fp, ">%sI ROT=0.0\n", s.name);
        row([&](int f) { return (resp[f].*(s.m)).value.imag() / s.unit; });
        std::fprintf(fp, ">%s.VAR\n", s.name);
        row([&](int f) {
            const double e = (resp[f].*(s.m)).error / s.unit;
            return e * e;
        });
        std::fprintf(fp, "\n");
    }
    std::fprintf(fp, ">END\n");
    if (std::fclose(fp) != 0) throw std::runtime_error("cannot write EDI file " + path);
}

void load_edi_directory(const std::string& dir, data::MTData& out) {
    namespace fs = std::filesystem;
    if (!fs::is_directory(dir)) throw std::runtime_error("not a directory: " + dir);
    std::vector<fs::path> files;
    for (const auto& e : fs::directory_iterator(dir))
        if (upper(e.path().extension().string()) == ".EDI") files.push_back(e.path());
    std::sort(files.begin(), files.end());
    if (files.empty()) throw std::runtime_error("no .edi files in " + dir);
    std::vector<EDIRecord> recs;
    for (const auto& f : files) {
        try {
            recs.push_back(read_edi(f.string()));
        } catch (const std::exception&) {
            // an unreadable file is skipped, as in the reference
        }
    }
    if (recs.empty()) throw std::runtime_error("no readable .edi files in " + dir);
    auto same = [](double a, double b) { return std::abs(a - b) / std::max(std::abs(a), 1e-30) < 1e-6; };
    std::vector<double> all;
    for (const auto& r : recs)
        for (double f : r.frequencies)
            if (std::none_of(all.begin(), all.end(), [&](double g) { return same(f, g); })) all.push_back(f);
    std::sort(all.begin(), all.end(), std::greater<double>());
    out = data::MTData();
    out.set_frequencies(all);
    for (const auto& r : recs) out.add_station(r.station);   // every station before any data
    data::MTResponse empty;   // all weights 0
    for (size_t s = 0; s < recs.size(); ++s) {
        for (size_t g = 0; g < all.size(); ++g) {
            const auto& fr = recs[s].frequencies;
            const auto it = std::find_if(fr.begin(), fr.end(), [&](double f) { return same(f, all[g]); });
            out.set_observed(static_cast<int>(s), static_cast<int>(g),
                             it == fr.end() ? empty : recs[s].responses[it - fr.begin()]);
        }
    }
}

void save_edi_directory(const std::string& dir, const data::MTData& d) {
    namespace fs = std::filesystem;
    fs::create_directories(dir);
    for (int s = 0; s < d.num_stations(); ++s) {
        const data::Station& st = d.station(s);
        const std::string name = st.name.empty() ? "station_" + std::to_string(s) : st.name;
        std::vector<data::MTResponse> r(d.num_frequencies());
        for (int f = 0; f < d.num_frequencies(); ++f) r[f] = d.observed(s, f);
        write_edi((fs::path(dir) / (name + ".edi")).string(), st, d.frequencies(), r);
    }
}

}  // namespace io
}  // namespace m3d
