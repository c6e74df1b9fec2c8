// Minimal MT data container with the API the forward solver and the inversion
// gradient use.  The reference's include/maple3dmt/data/mt_data.h is absent
// from its tree; this type provides the members its call sites need:
// frequencies(), num_stations(), station(i){name,x,y,z}, observed/predicted
// (s, f) -> MTResponse{Zxx,Zxy,Zyx,Zyy} of Datum{value,error,weight}
// (call sites: src/forward/forward_solver_fv.cpp:1720-1725,
// src/inversion/inversion_fv.cpp:258-271).
#pragma once

#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

namespace m3d {
namespace data {

struct Datum {
    std::complex<double> value{0.0, 0.0};
    double error = 0.0;
    double weight = 0.0;
};

struct MTResponse {
    Datum Zxx, Zxy, Zyx, Zyy, Tx, Ty;
};

struct Station {
    std::string name;
    double x = 0, y = 0, z = 0, lat = 0, lon = 0;
    bool has_geo = false;  // lat / lon given (EDI HEAD)
};

class MTData {
public:
    void add_station(const Station& s) {
        stations_.push_back(s);
        resize_();
    }
    void set_frequencies(const std::vector<double>& f) {
        freqs_ = f;
        resize_();
    }
    const std::vector<double>& frequencies() const { return freqs_; }
    int num_frequencies() const { return static_cast<int>(freqs_.size()); }
    int num_stations() const { return static_cast<int>(stations_.size()); }
    const Station& station(int i) const { return stations_.at(i); }

    const MTResponse& observed(int s, int f) const { return obs_.at(idx_(s, f)); }
    const MTResponse& predicted(int s, int f) const { return pred_.at(idx_(s, f)); }
    void set_observed(int s, int f, const MTResponse& r) { obs_.at(idx_(s, f)) = r; }
    void set_predicted(int s, int f, const MTResponse& r) { pred_.at(idx_(s, f)) = r; }

private:
    std::vector<Station> stations_;
    std::vector<double> freqs_;
    std::vector<MTResponse> obs_, pred_;
    size_t idx_(int s, int f) const {
        if (s < 0 || f < 0 || s >= num_stations() || f >= num_frequencies())
            throw std::out_of_range("MTData index");
        return size_t(f) * stations_.size() + size_t(s);
    }
    void resize_() {
        obs_.resize(stations_.size() * freqs_.size());
        pred_.resize(stations_.size() * freqs_.size());
    }
};

}  // namespace data
}  // namespace m3d
