// Exception hierarchy of the library: same class names, namespaces and
// meaning as the reference (include/lbm/exceptions/exceptions.hpp:31-155),
// each a std::runtime_error carrying the throw site.
#ifndef LBM_EXCEPTIONS_HPP
#define LBM_EXCEPTIONS_HPP

#include <source_location>
#include <stdexcept>
#include <string>
#include <string_view>

namespace lbm
{
namespace exceptions
{

class Exception : public std::runtime_error
{
  public:
    explicit Exception(const std::string &message,
                       const std::source_location &location = std::source_location::current())
        : Exception(message, "LBM Exception", location)
    {
    }

    // Name, message and throw site on several lines.
    std::string to_string() const;

    std::string_view name() const { return exception_name_; }
    const std::source_location &location() const { return source_location_; }

  protected:
    Exception(const std::string &message, std::string_view exception_name, const std::source_location &location)
        : std::runtime_error(message), exception_name_(exception_name), source_location_(location)
    {
    }

  private:
    std::string_view exception_name_;
    std::source_location source_location_;
};

namespace json
{
// A property has a value outside its valid set.
class PropertyArgumentException : public Exception
{
  public:
    explicit PropertyArgumentException(const std::string &message,
                                       const std::source_location &location = std::source_location::current())
        : Exception(message, "JSON Property Argument Exception", location)
    {
    }
};

// A required key is absent from the settings file.
class MissingPropertyException : public Exception
{
  public:
    explicit MissingPropertyException(const std::string &message,
                                      const std::source_location &location = std::source_location::current())
        : Exception(message, "JSON Missing Property Exception", location)
    {
    }
};

// A key of the settings file is not known or has the wrong type.
class UnknownPropertyException : public Exception
{
  public:
    explicit UnknownPropertyException(const std::string &message,
                                      const std::source_location &location = std::source_location::current())
        : Exception(message, "JSON Unknown Property Exception", location)
    {
    }
};
}  // namespace json

namespace observables
{
// A macroscopic observable left its valid range.
class OutOfBoundsException : public Exception
{
  public:
    explicit OutOfBoundsException(const std::string &message,
                                  const std::source_location &location = std::source_location::current())
        : Exception(message, "Observables Out Of Bounds Exception", location)
    {
    }
};
}  // namespace observables

namespace domain
{
// A node outside the simulation domain was addressed.
class OutOfDomainException : public Exception
{
  public:
    explicit OutOfDomainException(const std::string &message,
                                  const std::source_location &location = std::source_location::current())
        : Exception(message, "Out Of Domain Exception", location)
    {
    }
};
}  // namespace domain

namespace algorithm
{
// Waiting on a run that was never started, or whose worker failed.
class WaitException : public Exception
{
  public:
    explicit WaitException(const std::string &message,
                           const std::source_location &location = std::source_location::current())
        : Exception(message, "Wait Exception", location)
    {
    }
};
}  // namespace algorithm

}  // namespace exceptions
}  // namespace lbm

#endif  // LBM_EXCEPTIONS_HPP
