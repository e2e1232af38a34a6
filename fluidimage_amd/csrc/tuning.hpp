// Launch choices shared between the translation units of libpiv_hip: the
// TPS Cholesky's threads per subdomain for launches of more than two
// subdomains per compute unit (0: 256).
#pragma once

namespace fia {

int tps_chol_threads();
void set_tps_chol_threads(int threads);
int tps_shared_solve();
void set_tps_shared_solve(int mode);

}  // namespace fia
