// Fused single-pass MPDATA step for gfx950 (work in progress).
#include "mpdata_fused.hpp"

namespace mpdata {

template <typename T>
bool Fused<T>::supports(const Geo&, int, int, int) { return false; }
template <typename T>
const char* Fused<T>::scope() { return "(none yet)"; }
template <typename T>
hipError_t Fused<T>::step(const Geo&, T*, const T*, const T*, const T*, const T*, const T*,
                          const T*, const T*, hipStream_t) {
    return hipErrorNotSupported;
}
template struct Fused<double>;
template struct Fused<float>;

}  // namespace mpdata
