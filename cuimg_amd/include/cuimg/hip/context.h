// Host C++ side of the HIP tracking path: device context and error handling.
//
// Every call into the C ABI (include/cuimg_hip.h) goes through check(), which
// rethrows failures as std::runtime_error, as the reference's
// check_cuda_error does (cuimg/error.h:17-28).
#ifndef CUIMG_HIP_CONTEXT_H_
# define CUIMG_HIP_CONTEXT_H_

# include <memory>
# include <stdexcept>
# include <string>

# include <cuimg_hip.h>

namespace cuimg
{
  class hip_context
  {
  public:
    explicit hip_context(int device = 0)
    {
      cuimg_hip_ctx* c = nullptr;
      int rc = cuimg_hip_ctx_create(device, &c);
      if (rc)
        throw std::runtime_error("cuimg: no HIP device " + std::to_string(device) +
                                 " (error " + std::to_string(rc) + ")");
      ctx_ = std::shared_ptr<cuimg_hip_ctx>(c, cuimg_hip_ctx_destroy);
    }

    cuimg_hip_ctx* get() const { return ctx_.get(); }

    void check(int rc, const char* what) const
    {
      if (rc)
        throw std::runtime_error(std::string("cuimg: ") + what + ": " + cuimg_hip_last_error(ctx_.get()));
    }

    void sync() const { check(cuimg_hip_sync(get()), "sync"); }

    // Process-wide context on device 0, created on first use.
    static hip_context& default_context()
    {
      static hip_context c(0);
      return c;
    }

  private:
    std::shared_ptr<cuimg_hip_ctx> ctx_;
  };
}

#endif
