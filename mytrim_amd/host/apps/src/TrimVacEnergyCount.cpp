// TrimVacEnergyCount (apps/src/TrimVacEnergyCount.C:31-81): vacancies per
// (int ln E of the recoil, clamped at 0) and truncated 1 A depth bin.
#include "TrimVacEnergyCount.h"

#include <algorithm>
#include <cmath>
#include <fstream>

namespace MyTRIM_NS {

void TrimVacEnergyCount::vacancyCreation() {
  _simconf->vacancies_created++;
  const int x = int(_recoil->_pos(0));
  if (x < 0) return;
  int E = int(std::log(_recoil->_E));
  E = E < 0 ? 0 : E;
  if (E >= int(_evac_bin.size())) _evac_bin.resize(size_t(E) + 1);
  std::vector<unsigned int>& row = _evac_bin[size_t(E)];
  if (x >= int(row.size())) row.resize(size_t(x) + 1, 0);
  row[size_t(x)]++;
}

void TrimVacEnergyCount::threadJoin(const ThreadedTrimBase& other) {
  const TrimVacEnergyCount& o = static_cast<const TrimVacEnergyCount&>(other);
  _evac_bin.resize(std::max(_evac_bin.size(), o._evac_bin.size()));
  for (size_t E = 0; E < o._evac_bin.size(); ++E) {
    _evac_bin[E].resize(std::max(_evac_bin[E].size(), o._evac_bin[E].size()));
    for (size_t x = 0; x < o._evac_bin[E].size(); ++x) _evac_bin[E][x] += o._evac_bin[E][x];
  }
}

void TrimVacEnergyCount::writeOutput() {
  std::ofstream out(_base_name + "_evac.dat");
  for (size_t E = 0; E < _evac_bin.size(); ++E) {
    for (size_t x = 0; x < _evac_bin[E].size(); ++x) out << E << ' ' << x << ' ' << _evac_bin[E][x] << '\n';
    out << '\n';
  }
}

}  // namespace MyTRIM_NS
