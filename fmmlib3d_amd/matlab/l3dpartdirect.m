function [U]=l3dpartdirect(nsource,source,ifcharge,charge,ifdipole,dipstr,dipvec,ifpot,iffld,ntarget,target,ifpottarg,iffldtarg)
%L3DPARTDIRECT Laplace interactions in R^3 by direct summation on the MI355X.
%
% Same calling sequence and output struct as the reference's L3DPARTDIRECT
% (matlab/l3dpartdirect.m of fmmlib3d).
if nargin == 7, ifpot = 1; iffld = 1; ntarget = 0; target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
if nargin == 9, ntarget = 0; target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
if nargin == 11, ifpottarg = 1; iffldtarg = 1; end
if ntarget == 0, target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
[src, q, ds, dv, trg] = fmm3d_mi355x_args(nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, ntarget, target);
[pot, fld, pottarg, fldtarg] = fmm3d_mi355x_mex('l3dpartdirect', src, q, ds, dv, ...
    double(ifpot), double(iffld), trg, double(ifpottarg), double(iffldtarg));
if ifpot == 1, U.pot = pot; end
if iffld == 1, U.fld = fld; end
if ifpottarg == 1, U.pottarg = pottarg; end
if iffldtarg == 1, U.fldtarg = fldtarg; end
