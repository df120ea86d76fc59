// weldform/Material.h -- material classes with the constructors of the
// reference's Source/Material.h:64-151,202-220 (Elastic_, Material_,
// JohnsonCook, Hollomon), holding the parameters the device path needs.
// Yield and tangent laws themselves run on the device.
#pragma once
#include <cmath>
#include <string>

#include "matvec.h"

class Elastic_ {
 public:
  Elastic_() {}
  Elastic_(const double &e, const double &nu) : E_m(e), nu_m(nu) {}
  const double &E() const { return E_m; }
  const double &nu() const { return nu_m; }

 private:
  double E_m = 0.0, nu_m = 0.0;
};

class Material_;

// damage models (Source/Material.h:14-62): the Johnson-Cook criterion runs
// on the device path (Domain::model_damage); Rankine's pair criterion reads
// an unset strain threshold in the reference (Source/Damage.cpp:17,65) and
// is refused by the solver
class DamageModel {
 public:
  DamageModel() {}
  DamageModel(const double &smax_, const double &Gf_) : Gf(Gf_), sigma_max(smax_) {}
  virtual ~DamageModel() {}
  virtual std::string getDamageCriterion() { return std::string(""); }
  double Gf = 0.0, sigma_max = 0.0, delta_max = 0.0;
  Material_ *mat = nullptr;
};

class RankineDamage : public DamageModel {
 public:
  RankineDamage(const double &smax_, const double &Gf_) : DamageModel(smax_, Gf_) {}
  std::string getDamageCriterion() override { return std::string("Rankine"); }
};

// eps_f = (D1 + D2 exp(D3 s)) (1 + D4 ln(rate / eps_0)) (1 + D5 T_h), on the
// device (Source/Material.cpp:76-92)
class JohnsonCookDamage : public DamageModel {
 public:
  JohnsonCookDamage(const double &D1, const double &D2, const double &D3, const double &D4, const double &D5,
                    const double &eps0_)
      : m_D1(D1), m_D2(D2), m_D3(D3), m_D4(D4), m_D5(D5), m_eps0(eps0_) {}
  std::string getDamageCriterion() override { return std::string("JohnsonCook"); }
  double m_D1, m_D2, m_D3, m_D4, m_D5;
  double m_eps0;  // the material's reference strain rate (the device takes the material's)
};

class Material_ {
 public:
  enum Kind { Plain = 0, KindHollomon = 1, KindJohnsonCook = 2 };
  Material_() {}
  explicit Material_(const Elastic_ el) : elastic_m(el) {}
  virtual ~Material_() {}
  const Elastic_ &Elastic() const { return elastic_m; }
  virtual Kind kind() const { return Plain; }
  // yield stress at (strain, strain rate, temperature); the base class
  // returns 0 (Source/Material.h:98-101)
  virtual double CalcYieldStress(const double &, const double &, const double &) { return 0.0; }
  double &getT_t() { return T_t; }
  double &getT_m() { return T_m; }
  DamageModel *damage = nullptr;

 protected:
  Elastic_ elastic_m;
  double T_t = 0.0, T_m = 0.0;
};

// sy = (A + B eps_p^n) (1 + C ln(eps_dot / eps_0)) (1 - T_h^m)
// argument order of Source/Material.h:128-136: A, B, n, C, eps_0, m, T_m, T_t
class JohnsonCook : public Material_ {
 public:
  JohnsonCook() {}
  JohnsonCook(const Elastic_ &el, const double &a, const double &b, const double &n_, const double &c,
              const double &eps_0_, const double &m_, const double &T_m_, const double &T_t_)
      : Material_(el), A(a), B(b), C(c), n(n_), m(m_), eps_0(eps_0_) {
    T_m = T_m_;
    T_t = T_t_;
  }
  Kind kind() const override { return KindJohnsonCook; }
  double &getRefStrainRate() { return eps_0; }
  // Source/Material.cpp:6-25
  double CalcYieldStress(const double &strain, const double &strain_rate, const double &temp) override {
    const double T_h = (temp - T_t) / (T_m - T_t);
    double f = 1.0;
    if (strain_rate > eps_0) f = (1.0 + C * std::log(strain_rate / eps_0));
    return (A + B * std::pow(strain, n)) * f * (1.0 - std::pow(T_h, m));
  }
  double A = 0, B = 0, C = 0, n = 0, m = 0, eps_0 = 1.0;
};

// power law with plateau, Source/Material.cpp:48-57
class Hollomon : public Material_ {
 public:
  Hollomon() {}
  Hollomon(const Elastic_ &el, const double sy0_, const double &k_, const double &m_)
      : Material_(el), K(k_), m(m_), sy0(sy0_) {
    eps0 = sy0_ / el.E();
    eps1 = std::pow(sy0_ / k_, 1. / m_);
    if (eps0 > eps1)
      throw new Fatal(
          "ERROR, Hollomon material bad definition, please correct Yield Stress, Elastic Modulus or Material "
          "hardening constants.");
  }
  Kind kind() const override { return KindHollomon; }
  // Source/Material.cpp:59-64.  Hollomon overrides only this one-argument
  // form, so the three-argument call of the JSON driver reaches the base
  // class and gives 0 (Source/Material.h:204-224), as in the reference.
  using Material_::CalcYieldStress;
  double CalcYieldStress(const double &strain) {
    if (strain + eps0 > eps1) return K * std::pow(strain + eps0, m);
    return sy0;
  }
  double K = 0, m = 1, eps0 = 0, eps1 = 0, sy0 = 0;
};
