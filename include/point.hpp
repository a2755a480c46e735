// point.hpp -- a data point with its cluster label.
//
// Same interface as the reference's Point<T> (src/include/point.hpp:11-81):
// construct from a value array, get(i), getCluster/setCluster, operator== by
// value, stream output "cluster x0 x1 ...". Kmeans::run writes the labels
// back through setCluster (src/kmeans.cu:1101-1103).
#ifndef KMC_POINT_HPP
#define KMC_POINT_HPP

#include <cstdlib>
#include <iomanip>
#include <ostream>
#include <vector>

template <typename T>
class Point {
 public:
  Point() = default;
  Point(const T* x, unsigned int d) : x_(x, x + d) {}
  explicit Point(const Point<T>* p) : x_(p->x_), cluster_(p->cluster_) {}

  int getCluster() const { return static_cast<int>(cluster_); }
  void setCluster(unsigned int c) { cluster_ = c; }

  // Out-of-range access ends the program with exit code 3, the reference's
  // EXIT_POINT_IOB (src/include/utils.hpp:36).
  T get(unsigned int i) const {
    if (i >= x_.size()) std::exit(3);
    return x_[i];
  }
  unsigned int size() const { return static_cast<unsigned int>(x_.size()); }
  const T* data() const { return x_.data(); }

  bool operator==(const Point<T>& p) const {
    if (x_.size() != p.x_.size()) return false;
    for (size_t i = 0; i < x_.size(); ++i)
      if (x_[i] != p.x_[i]) return false;
    return true;
  }

  friend std::ostream& operator<<(std::ostream& os, const Point& p) {
    os << std::setw(9) << p.cluster_;
    for (const T& v : p.x_) os << std::setw(9) << std::setprecision(5) << v;
    return os;
  }

 private:
  std::vector<T> x_;
  unsigned int cluster_ = 0;
};

#endif  // KMC_POINT_HPP
