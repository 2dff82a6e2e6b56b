// Image containers of the r3d API: strided non-owning views and owning images
// with 64-byte aligned rows.  Same interface as the reference's
// src/r3d/core/image.hpp (ImageView, Image, ImageU8/U16/F32), so callers of the
// matcher compile unchanged.
#pragma once

#include <cassert>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <type_traits>

#include "r3d/core/error.hpp"

namespace r3d {

inline constexpr std::size_t kBufferAlignment = 64;

template <typename T>
class ImageView {
 public:
  ImageView() = default;
  ImageView(T* data, int width, int height, int channels, std::ptrdiff_t stride)
      : data_(data), width_(width), height_(height), channels_(channels), stride_(stride) {
    assert(width >= 0 && height >= 0 && channels > 0);
    assert(stride >= static_cast<std::ptrdiff_t>(width) * channels);
  }
  // ImageView<T> -> ImageView<const T>
  template <typename U, typename = std::enable_if_t<std::is_same_v<const U, T> && !std::is_same_v<U, T>>>
  ImageView(const ImageView<U>& o)  // NOLINT(google-explicit-constructor)
      : ImageView(o.data(), o.width(), o.height(), o.channels(), o.stride()) {}

  [[nodiscard]] T* data() const noexcept { return data_; }
  [[nodiscard]] int width() const noexcept { return width_; }
  [[nodiscard]] int height() const noexcept { return height_; }
  [[nodiscard]] int channels() const noexcept { return channels_; }
  [[nodiscard]] std::ptrdiff_t stride() const noexcept { return stride_; }
  [[nodiscard]] bool empty() const noexcept { return width_ == 0 || height_ == 0; }
  [[nodiscard]] T* row(int y) const noexcept { return data_ + static_cast<std::ptrdiff_t>(y) * stride_; }
  [[nodiscard]] T& operator()(int x, int y, int c = 0) const noexcept {
    return row(y)[static_cast<std::ptrdiff_t>(x) * channels_ + c];
  }
  [[nodiscard]] ImageView subview(int x, int y, int w, int h) const {
    return ImageView(row(y) + static_cast<std::ptrdiff_t>(x) * channels_, w, h, channels_, stride_);
  }

 private:
  T* data_ = nullptr;
  int width_ = 0;
  int height_ = 0;
  int channels_ = 1;
  std::ptrdiff_t stride_ = 0;
};

template <typename T>
class Image {
  static_assert(std::is_trivially_copyable_v<T>, "Image requires a trivially copyable pixel type");

 public:
  Image() = default;
  Image(int width, int height, int channels = 1) : Image(uninitialized(width, height, channels)) { fill(T{}); }
  Image(const Image& o) : Image(uninitialized(o.width_, o.height_, o.channels_)) {
    if (size_ > 0) std::memcpy(data_.get(), o.data_.get(), size_ * sizeof(T));
  }
  Image& operator=(const Image& o) {
    if (this != &o) *this = Image(o);
    return *this;
  }
  Image(Image&&) noexcept = default;
  Image& operator=(Image&&) noexcept = default;

  [[nodiscard]] static Image uninitialized(int width, int height, int channels = 1) {
    if (width < 0 || height < 0 || channels <= 0) throw Error("invalid image dimensions");
    Image img;
    img.width_ = width;
    img.height_ = height;
    img.channels_ = channels;
    img.stride_ = aligned_stride(width, channels);
    img.size_ = static_cast<std::size_t>(img.stride_) * static_cast<std::size_t>(height);
    if (img.size_ > 0) {
      void* p = std::aligned_alloc(kBufferAlignment,
                                   (img.size_ * sizeof(T) + kBufferAlignment - 1) / kBufferAlignment * kBufferAlignment);
      if (p == nullptr) throw std::bad_alloc();
      img.data_.reset(static_cast<T*>(p));
    }
    return img;
  }

  [[nodiscard]] int width() const noexcept { return width_; }
  [[nodiscard]] int height() const noexcept { return height_; }
  [[nodiscard]] int channels() const noexcept { return channels_; }
  [[nodiscard]] std::ptrdiff_t stride() const noexcept { return stride_; }
  [[nodiscard]] bool empty() const noexcept { return width_ == 0 || height_ == 0; }
  [[nodiscard]] T* data() noexcept { return data_.get(); }
  [[nodiscard]] const T* data() const noexcept { return data_.get(); }
  [[nodiscard]] T* row(int y) noexcept { return data_.get() + static_cast<std::ptrdiff_t>(y) * stride_; }
  [[nodiscard]] const T* row(int y) const noexcept { return data_.get() + static_cast<std::ptrdiff_t>(y) * stride_; }
  [[nodiscard]] T& operator()(int x, int y, int c = 0) noexcept {
    return row(y)[static_cast<std::ptrdiff_t>(x) * channels_ + c];
  }
  [[nodiscard]] const T& operator()(int x, int y, int c = 0) const noexcept {
    return row(y)[static_cast<std::ptrdiff_t>(x) * channels_ + c];
  }
  [[nodiscard]] ImageView<T> view() noexcept { return {data(), width_, height_, channels_, stride_}; }
  [[nodiscard]] ImageView<const T> view() const noexcept { return {data(), width_, height_, channels_, stride_}; }
  void fill(const T& v) {
    for (std::size_t i = 0; i < size_; ++i) data_.get()[i] = v;
  }
  [[nodiscard]] static std::ptrdiff_t aligned_stride(int width, int channels) noexcept {
    constexpr auto per_line = static_cast<std::ptrdiff_t>(kBufferAlignment / sizeof(T));
    const std::ptrdiff_t e = static_cast<std::ptrdiff_t>(width) * channels;
    return (e + per_line - 1) / per_line * per_line;
  }

 private:
  struct Free {
    void operator()(T* p) const noexcept { std::free(p); }
  };
  std::unique_ptr<T, Free> data_;
  std::size_t size_ = 0;
  int width_ = 0;
  int height_ = 0;
  int channels_ = 1;
  std::ptrdiff_t stride_ = 0;
};

using ImageU8 = Image<std::uint8_t>;
using ImageU16 = Image<std::uint16_t>;
using ImageF32 = Image<float>;

}  // namespace r3d
