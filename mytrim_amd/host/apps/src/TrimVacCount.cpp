// TrimVacCount (apps/src/TrimVacCount.C:31-79): the reference's truncating
// depth bins and output format.
#include "TrimVacCount.h"

#include <algorithm>
#include <fstream>

namespace MyTRIM_NS {

void TrimVacCount::count(std::vector<unsigned int>& bins, Real x_real) {
  const int x = int(x_real);  // truncation toward zero: (-1, 0) lands in bin 0
  if (x < 0) return;
  if (x >= int(bins.size())) bins.resize(size_t(x) + 1, 0);
  bins[size_t(x)]++;
}

void TrimVacCount::vacancyCreation() {
  _simconf->vacancies_created++;
  count(_vac_bin, _recoil->_pos(0));
}

void TrimVacCount::replacementCollision() { count(_repl_bin, _recoil->_pos(0)); }

void TrimVacCount::threadJoin(const ThreadedTrimBase& other) {
  const TrimVacCount& o = static_cast<const TrimVacCount&>(other);
  _vac_bin.resize(std::max(_vac_bin.size(), o._vac_bin.size()));
  for (size_t x = 0; x < o._vac_bin.size(); ++x) _vac_bin[x] += o._vac_bin[x];
  _repl_bin.resize(std::max(_repl_bin.size(), o._repl_bin.size()));
  for (size_t x = 0; x < o._repl_bin.size(); ++x) _repl_bin[x] += o._repl_bin[x];
}

void TrimVacCount::writeOutput() {
  const size_t n = std::max(_vac_bin.size(), _repl_bin.size());
  _vac_bin.resize(n);
  _repl_bin.resize(n);
  std::ofstream out(_base_name + "_vac.dat");
  for (size_t x = 0; x < n; ++x) out << x << ' ' << _vac_bin[x] << ' ' << _repl_bin[x] << '\n';
}

}  // namespace MyTRIM_NS
