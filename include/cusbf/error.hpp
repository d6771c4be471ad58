// cusbf/error.hpp -- Result<T> / Error surface of the C++ API (host only).
//
// Mirrors include/cusbf/error.hpp of the reference: five error categories with
// pinned indices (error.hpp:67-73, 139-143), Result<T> with ok()/err(),
// Err()/Ok(), CUSBF_TRY / CUSBF_UNWRAP (GNU statement expressions,
// error.hpp:246-263). libstdc++ 11 has no std::expected, so Result is a small
// variant of its own with the same members callers use (operator bool, *, ->,
// value(), error()).
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <variant>

namespace cusbf {

enum class ErrorCategory : uint8_t {
    cuda,  // device runtime errors (HIP on this build)
    io,
    fastx_parse,
    invalid_argument,
    resource,
};
static_assert(static_cast<int>(ErrorCategory::cuda) == 0);
static_assert(static_cast<int>(ErrorCategory::io) == 1);
static_assert(static_cast<int>(ErrorCategory::fastx_parse) == 2);
static_assert(static_cast<int>(ErrorCategory::invalid_argument) == 3);
static_assert(static_cast<int>(ErrorCategory::resource) == 4);

/// error.hpp:20-41: file, line and column of a FASTX input position.
struct SourceLocation {
    std::string file;
    uint32_t line{};
    uint32_t column{};

    [[nodiscard]] std::string to_string() const {
        return file + ":" + std::to_string(line) + ":" + std::to_string(column);
    }
};

/// error.hpp:50-53: a FASTX parse error and where it happened.
struct FastxParseError {
    SourceLocation location;
    std::string message;
};

class Error {
   public:
    Error(ErrorCategory category, std::string message)
        : category_(category), message_(std::move(message)) {}

    [[nodiscard]] ErrorCategory category() const noexcept { return category_; }
    [[nodiscard]] const std::string& message() const noexcept { return message_; }

    /// error.hpp:97-99: the parse error payload (nullptr for other categories).
    [[nodiscard]] const FastxParseError* as_fastx_parse() const noexcept {
        return category_ == ErrorCategory::fastx_parse ? &parse_ : nullptr;
    }

    /// error.hpp:123-129: message "file:line:column: detail".
    [[nodiscard]] static Error fastx_parse(SourceLocation site, std::string message) {
        Error e(ErrorCategory::fastx_parse, message);
        e.parse_ = FastxParseError{std::move(site), std::move(message)};
        return e;
    }

    [[nodiscard]] static Error cuda(std::string message) {
        return Error(ErrorCategory::cuda, std::move(message));
    }
    [[nodiscard]] static Error io(std::string message) {
        return Error(ErrorCategory::io, std::move(message));
    }
    [[nodiscard]] static Error invalid_argument(std::string message) {
        return Error(ErrorCategory::invalid_argument, std::move(message));
    }
    [[nodiscard]] static Error resource(std::string message) {
        return Error(ErrorCategory::resource, std::move(message));
    }

   private:
    ErrorCategory category_;
    std::string message_;
    FastxParseError parse_;
};

struct Unexpected {
    Error error;
};

/// Failure return; converts to any Result<T>.
[[nodiscard]] inline Unexpected Err(Error error) { return Unexpected{std::move(error)}; }

class bad_result_access : public std::runtime_error {
   public:
    explicit bad_result_access(const Error& e) : std::runtime_error(e.message()) {}
};

template <typename T>
class [[nodiscard]] Result {
   public:
    Result(T value) : v_(std::in_place_index<0>, std::move(value)) {}  // NOLINT: implicit like expected
    Result(Unexpected e) : v_(std::in_place_index<1>, std::move(e.error)) {}  // NOLINT

    [[nodiscard]] static Result ok(T value) { return Result(std::move(value)); }
    [[nodiscard]] static Result err(Error error) { return Result(Unexpected{std::move(error)}); }

    [[nodiscard]] bool has_value() const noexcept { return v_.index() == 0; }
    explicit operator bool() const noexcept { return has_value(); }
    T& operator*() { return std::get<0>(v_); }
    const T& operator*() const { return std::get<0>(v_); }
    T* operator->() { return &std::get<0>(v_); }
    const T* operator->() const { return &std::get<0>(v_); }
    T& value() {
        if (!has_value()) throw bad_result_access(error());
        return std::get<0>(v_);
    }
    const T& value() const {
        if (!has_value()) throw bad_result_access(error());
        return std::get<0>(v_);
    }
    [[nodiscard]] const Error& error() const { return std::get<1>(v_); }

   private:
    std::variant<T, Error> v_;
};

template <>
class [[nodiscard]] Result<void> {
   public:
    Result() = default;
    Result(Unexpected e) : error_(std::move(e.error)), failed_(true) {}  // NOLINT

    [[nodiscard]] static Result ok() { return {}; }
    [[nodiscard]] static Result err(Error error) { return Result(Unexpected{std::move(error)}); }

    [[nodiscard]] bool has_value() const noexcept { return !failed_; }
    explicit operator bool() const noexcept { return has_value(); }
    void value() const {
        if (failed_) throw bad_result_access(error_);
    }
    [[nodiscard]] const Error& error() const { return error_; }

   private:
    Error error_{ErrorCategory::cuda, std::string()};
    bool failed_ = false;
};

/// Success return for Result<void>.
[[nodiscard]] inline Result<void> Ok() noexcept { return Result<void>::ok(); }

/// Aborts when a Result<void> failed (benchmarks and tests), error.hpp:178-184.
inline void require_void(const Result<void>& result) {
    if (!result) {
        std::fputs(result.error().message().c_str(), stderr);
        std::fputc('\n', stderr);
        std::abort();
    }
}

namespace detail {

template <typename T>
[[nodiscard]] T try_unwrap_success(Result<T>& result) {
    return std::move(*result);
}
inline void try_unwrap_success(Result<void>&) {}
[[nodiscard]] inline Unexpected propagate_error(const Error& error) { return Unexpected{error}; }

}  // namespace detail

}  // namespace cusbf

/// Propagates a failure from the enclosing function; yields the value on success.
#define CUSBF_TRY(expr)                                                     \
    ({                                                                      \
        auto _cusbf_result = (expr);                                        \
        if (!_cusbf_result) {                                               \
            return ::cusbf::detail::propagate_error(_cusbf_result.error()); \
        }                                                                   \
        ::cusbf::detail::try_unwrap_success(_cusbf_result);                 \
    })

/// Unwraps a Result or throws std::runtime_error (tests and applications).
#define CUSBF_UNWRAP(expr)                                             \
    ({                                                                 \
        auto _cusbf_result = (expr);                                   \
        if (!_cusbf_result) {                                          \
            throw std::runtime_error(_cusbf_result.error().message()); \
        }                                                              \
        ::cusbf::detail::try_unwrap_success(_cusbf_result);            \
    })
