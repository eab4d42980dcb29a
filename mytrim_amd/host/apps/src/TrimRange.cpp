// TrimRange (apps/src/TrimRange.C:31-121).
#include "TrimRange.h"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <utility>

namespace MyTRIM_NS {

// NRT damage of the recoil (TrimRange.C:36-52); the int counter takes the
// fractional estimate by truncating the running sum
void TrimRange::vacancyCreation() {
  const Real Ed = _element->_Edisp;
  const Real ed = 0.0115 * std::pow(_recoil->_Z, -7.0 / 3.0) * _recoil->_E;
  const Real kd = 0.1337 * std::pow(_recoil->_Z, 2.0 / 3.0) / std::sqrt(_recoil->_m);
  const Real g = 3.4008 * std::pow(ed, 1.0 / 6.0) + 0.40244 * std::pow(ed, 3.0 / 4.0) + ed;
  const Real Ev = _recoil->_E / (1.0 + kd * g);
  if (Ev < Ed) return;
  if (Ev >= Ed / 0.4)
    _simconf->vacancies_created += Ev * 0.4 / Ed;
  else
    _simconf->vacancies_created++;
}

void TrimRange::dissipateRecoilEnergy() { _range[size_t(_recoil->_Z)].push_back(_recoil->_pos(0)); }

void TrimRange::threadJoin(const ThreadedTrimBase& other) {
  const TrimRange& o = static_cast<const TrimRange&>(other);
  for (size_t Z = 0; Z < _range.size(); ++Z) _range[Z].insert(_range[Z].end(), o._range[Z].begin(), o._range[Z].end());
}

void TrimRange::writeOutput() {
  unsigned int max_samples = 0;
  Real x_min = 0.0, x_max = 0.0;
  for (const auto& r : _range) {
    if (r.size() > max_samples) max_samples = unsigned(r.size());
    for (Real x : r) {
      x_min = std::min(x_min, x);
      x_max = std::max(x_max, x);
    }
  }
  std::ofstream out(_base_name + "_ranges.dat");
  out << "#x";
  if (max_samples == 0 || !(x_max > x_min)) {  // nothing came to rest: header only
    out << '\n';
    return;
  }
  // bin width heuristic of the reference
  const Real xwidth = x_max - x_min;
  const Real bin = std::min(xwidth * 100.0 / max_samples, xwidth / 10.0);
  const unsigned int nbin = xwidth / bin + 1;
  std::vector<std::pair<int, std::vector<unsigned int>>> hists;
  for (size_t Z = 0; Z < _range.size(); ++Z) {
    if (_range[Z].empty()) continue;
    std::vector<unsigned int> h(nbin);
    for (Real x : _range[Z]) {
      const size_t b = size_t(std::floor((x - x_min) / bin));
      if (b < h.size()) h[b]++;
    }
    hists.emplace_back(int(Z), std::move(h));
  }
  for (const auto& h : hists) out << " Z" << h.first;
  out << '\n';
  for (unsigned int n = 0; n < nbin; ++n) {
    out << (n * bin + x_min);
    for (const auto& h : hists) out << ' ' << h.second[n];
    out << '\n';
  }
}

}  // namespace MyTRIM_NS
