/* Test fixture: the smallest Complex<T> with the layout of the caller's
 * framework/lib/Complex.h (a class template holding T re, im in that order), so the
 * header-only shims in include/ compile without the caller's tree. */
#ifndef COMPLEX_H
#define COMPLEX_H
template <class T>
class Complex {
 public:
  T re, im;
  Complex() : re(0), im(0) {}
  Complex(T r, T i) : re(r), im(i) {}
};
#endif
