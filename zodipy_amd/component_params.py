"""Geometric parameter tables of the built-in models.

The values are the published model parameters and must equal the reference's
tables bit for bit (zodipy/component_params.py:19-206): DIRBE (Kelsall et al.
1998, six components), PLANCK (DIRBE without ring and feature) and the
experimental Rowan-Robinson & May (2013) model.
"""

from __future__ import annotations

from zodipy_amd.component import (
    Band,
    BroadBand,
    Cloud,
    Comet,
    ComponentLabel as L,
    Fan,
    Feature,
    FeatureRRM,
    Interstellar,
    NarrowBand,
    Ring,
    RingRRM,
    ZodiacalComponent,
)

DIRBE: dict[L, ZodiacalComponent] = {
    L.CLOUD: Cloud(
        x_0=0.011887800744346281, y_0=0.0054765064662263777, z_0=-0.0021530908020710744,
        i=2.0335188072390769, Omega=77.657955554097114,
        n_0=1.1344373881427960e-07, alpha=1.3370696705930281, beta=4.1415004157586637,
        gamma=0.94206179393358036, mu=0.18873176489090190,
    ),
    L.BAND1: Band(
        x_0=0.0, y_0=0.0, z_0=0.0, i=0.56438265154389733, Omega=80.0,
        n_0=5.5890290403228370e-10, delta_zeta=8.7850534408713035,
        v=0.10000000149011612, p=4.0, delta_r=1.5,
    ),
    L.BAND2: Band(
        x_0=0.0, y_0=0.0, z_0=0.0, i=1.2000000476837158, Omega=30.347475578624532,
        n_0=1.9877609422590801e-09, delta_zeta=1.9917032425777641,
        v=0.89999997615814209, p=4.0, delta_r=0.94121881201651147,
    ),
    L.BAND3: Band(
        x_0=0.0, y_0=0.0, z_0=0.0, i=0.80000001192092896, Omega=80.0,
        n_0=1.4369827283512384e-10, delta_zeta=15,
        v=0.050000000745058060, p=4.0, delta_r=1.5,
    ),
    L.RING: Ring(
        x_0=0.0, y_0=0.0, z_0=0.0, i=0.48707166006819241, Omega=22.278979678854448,
        n_0=1.8260527826501675e-08, R=1.0281924326308751,
        sigma_r=0.025000000372529030, sigma_z=0.054068037356978099,
    ),
    L.FEATURE: Feature(
        x_0=0.0, y_0=0.0, z_0=0.0, i=0.48707166006819241, Omega=22.278979678854448,
        n_0=2.0094267183590947e-08, R=1.0579182694524214,
        sigma_r=0.10287314662396611, sigma_z=0.091442963768716023,
        theta=-10.0, sigma_theta=12.115210933938741,
    ),
}

PLANCK: dict[L, ZodiacalComponent] = {
    label: comp for label, comp in DIRBE.items() if label not in (L.RING, L.FEATURE)
}

R_MARS = 1.5237
R_ASTEROID_BELT = 3.137
R_KUIPER_BELT = 30

RRM: dict[L, ZodiacalComponent] = {
    L.FAN: Fan(
        x_0=0.0, y_0=0.0, z_0=0.0, i=1.5, Omega=78,
        gamma=1.3, Z_0=0.06, Q=10.7, P=2.13, R_outer=R_MARS,
    ),
    L.COMET: Comet(
        x_0=0.0, y_0=0.0, z_0=0.0, i=1.5, Omega=78,
        gamma=1, Z_0=0.06, P=2.5, amp=0.37, R_inner=R_MARS, R_outer=R_KUIPER_BELT,
    ),
    L.INNER_NARROW_BAND: NarrowBand(
        x_0=0.0, y_0=0.0, z_0=0.0, i=1.5, Omega=78,
        gamma=1, A=0.04, G=0.12, R_inner=R_MARS, R_outer=R_ASTEROID_BELT, beta_nb=9.35,
    ),
    L.OUTER_NARROW_BAND: NarrowBand(
        x_0=0.0, y_0=0.0, z_0=0.0, i=1.5, Omega=78,
        gamma=1, A=0.032, G=0.6, R_inner=R_MARS, R_outer=R_ASTEROID_BELT, beta_nb=1.42,
    ),
    L.BROAD_BAND: BroadBand(
        x_0=0.0, y_0=0.0, z_0=0.0, i=2.6, Omega=110,
        gamma=1, A=0.051, R_inner=R_MARS, R_outer=R_ASTEROID_BELT, beta_bb=9.3, sigma_bb=5,
    ),
    L.INTERSTELLAR: Interstellar(x_0=0.0, y_0=0.0, z_0=0.0, i=0, Omega=0, amp=0.010),
    L.RING_RRM: RingRRM(
        x_0=0.0, y_0=0.0, z_0=0.0, i=0.49, Omega=22.3,
        n_0=1, R=1.03, sigma_r=0.025, sigma_z=0.054, A=0.16,
    ),
    L.FEATURE_RRM: FeatureRRM(
        x_0=0.0, y_0=0.0, z_0=0.0, i=0.49, Omega=22.3,
        n_0=1, R=1.06, sigma_r=0.10, sigma_z=0.091, theta=-10.0, sigma_theta=12.1, A=0.065,
    ),
}
