// device_image2d<V>: device image in HBM behind the C ABI
// (cuimg/gpu/device_image2d.h:24-80, .hpp:23-38): shared ownership, shallow
// copies, 128-byte aligned rows; kernel() gives the non-owning view that the
// C ABI takes (the reference's kernel_image2d, cuimg/gpu/kernel_image2d.h).
#ifndef CUIMG_DEVICE_IMAGE2D_H_
# define CUIMG_DEVICE_IMAGE2D_H_

# include <memory>

# include <cuimg/cpu/host_image2d.h>
# include <cuimg/gl.h>
# include <cuimg/hip/context.h>

namespace cuimg
{
  template <typename V> struct hip_dtype;
  template <> struct hip_dtype<gl8u> { static const int value = CUIMG_U8; };
  template <> struct hip_dtype<unsigned char> { static const int value = CUIMG_U8; };
  template <> struct hip_dtype<float> { static const int value = CUIMG_F32; };
  template <> struct hip_dtype<unsigned int> { static const int value = CUIMG_U32; };

  template <typename V>
  class device_image2d
  {
  public:
    typedef V value_type;

    device_image2d() {}
    device_image2d(const obox2d& d, unsigned border = 0, hip_context& ctx = hip_context::default_context())
      : ctx_(&ctx)
    {
      auto* img = new cuimg_img();
      int rc = cuimg_hip_image_alloc(ctx.get(), d.nrows(), d.ncols(), border, hip_dtype<V>::value, img);
      if (rc)
      {
        delete img;
        ctx.check(rc, "image_alloc");
      }
      cuimg_hip_ctx* c = ctx.get();
      img_ = std::shared_ptr<cuimg_img>(img, [c](cuimg_img* p) {
        cuimg_hip_image_free(c, p);
        delete p;
      });
    }

    obox2d domain() const { return obox2d(img_->nrows, img_->ncols); }
    int nrows() const { return img_->nrows; }
    int ncols() const { return img_->ncols; }
    int border() const { return img_->border; }
    size_t pitch() const { return img_->pitch; }
    const cuimg_img& kernel() const { return *img_; }
    hip_context& context() const { return *ctx_; }

  private:
    hip_context* ctx_ = nullptr;
    std::shared_ptr<cuimg_img> img_;
  };

  // copy(host, device) / copy(device, host) of the domain (cuimg/copy.h:59-71).
  template <typename V>
  void copy(const host_image2d<V>& in, device_image2d<V>& out)
  {
    out.context().check(cuimg_hip_image_upload(out.context().get(), &out.kernel(), in.begin(), in.pitch()),
                        "copy(host, device)");
  }

  template <typename V>
  void copy(const device_image2d<V>& in, host_image2d<V>& out)
  {
    in.context().check(cuimg_hip_image_download(in.context().get(), &in.kernel(), out.begin(), out.pitch()),
                       "copy(device, host)");
  }
}

#endif
