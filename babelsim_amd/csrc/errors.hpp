// Error plumbing for the C ABI: exceptions inside, int codes + a thread-local
// message outside (the host adaptor turns them back into std::invalid_argument
// / std::runtime_error as the reference's seam expects).
#pragma once

#include <stdexcept>
#include <string>

namespace bsim {

void setError(const std::string& message);

template <class F> int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument& e) {
        setError(e.what());
        return 2;
    } catch (const std::exception& e) {
        setError(e.what());
        return 1;
    } catch (...) {
        setError("unknown error");
        return 1;
    }
}

}  // namespace bsim
