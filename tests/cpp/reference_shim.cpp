// TEST INFRASTRUCTURE: detail::sgm_cost_volume_reference (the reference's
// textbook volume, stereo_matcher.hpp:105-109) for C++ tests of the r3d API,
// provided by the CPU oracle.
#include <cstring>

#include "../../oracle/oracle.hpp"
#include "r3d/stereo/stereo_matcher.hpp"

namespace r3d::detail {

namespace {
template <typename T>
oracle::Img<T> copy(ImageView<const T> v) {
  oracle::Img<T> o(v.width(), v.height());
  for (int y = 0; y < v.height(); ++y) std::memcpy(o.row(y), v.row(y), sizeof(T) * static_cast<std::size_t>(v.width()));
  return o;
}
}  // namespace

std::vector<std::uint16_t> sgm_cost_volume_reference(ImageView<const std::uint16_t> left,
                                                     ImageView<const std::uint16_t> right, int disparity_min,
                                                     int num_disparities, const StereoOptions& options,
                                                     const LabelRanges* ranges) {
  r3dg_options o{};
  r3dg_default_options(&o);
  o.p1 = options.p1;
  o.p2 = options.p2;
  o.p2_adapt = options.p2_adapt;
  o.invalid_cost = options.invalid_cost;
  oracle::Ranges r;
  if (ranges != nullptr) r = oracle::Ranges{copy(ranges->lo.view()), copy(ranges->num.view())};
  return oracle::textbook_volume(copy(left), copy(right), disparity_min, num_disparities, o,
                                 ranges != nullptr ? &r : nullptr);
}

}  // namespace r3d::detail
