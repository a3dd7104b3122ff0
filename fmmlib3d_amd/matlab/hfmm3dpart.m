function [U]=hfmm3dpart(iprec,zk,nsource,source,ifcharge,charge,ifdipole,dipstr,dipvec,ifpot,iffld,ntarget,target,ifpottarg,iffldtarg)
%HFMM3DPART Helmholtz particle FMM in R^3 on the MI355X (fmm3d-mi355x).
%
% Same calling sequence and output struct as the reference's HFMM3DPART
% (matlab/hfmm3dpart.m of fmmlib3d); kernel e^{i zk r}/r, zk complex
% (zk = 0 returns ier = 128).  Arguments and outputs otherwise as
% LFMM3DPART.
if nargin == 9, ifpot = 1; iffld = 1; ntarget = 0; target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
if nargin == 11, ntarget = 0; target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
if nargin == 13, ifpottarg = 1; iffldtarg = 1; end
if ntarget == 0, target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
[src, q, ds, dv, trg] = fmm3d_mi355x_args(nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, ntarget, target);
[ier, pot, fld, pottarg, fldtarg] = fmm3d_mi355x_mex('hfmm3dparttarg', iprec, complex(zk), src, q, ds, dv, ...
    double(ifpot), double(iffld), trg, double(ifpottarg), double(iffldtarg));
if ifpot == 1, U.pot = pot; end
if iffld == 1, U.fld = fld; end
if ifpottarg == 1, U.pottarg = pottarg; end
if iffldtarg == 1, U.fldtarg = fldtarg; end
U.ier = ier;
